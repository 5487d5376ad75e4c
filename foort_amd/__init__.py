"""foort_amd -- MI355X (gfx950) per-pixel null-geodesic integrator with the
plugin surface of FOORT (Metric / Diagnostic / Termination / Mesh /
ViewScreen).  This module is the ctypes binding of the C ABI in
include/foort_gpu.h; the integration itself runs in libfoort_gpu.so.

There is no CPU fallback: if the library is missing, or no GPU is visible,
the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import RunDesc, GeodesicResults, Positions, Stats, make_desc, default_desc  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FOORT_GPU_LIB") or os.path.join(_HERE, "libfoort_gpu.so")
_lib = None

EXPORTS = [
    "foort_desc_init", "foort_desc_size", "foort_gpu_open", "foort_gpu_close", "foort_gpu_last_error",
    "foort_gpu_trace_grid", "foort_gpu_trace_pixels", "foort_gpu_trace_device", "foort_gpu_camera",
    "foort_gpu_device_alloc", "foort_gpu_device_free", "foort_gpu_copy_to_device", "foort_gpu_copy_to_host",
    "foort_gpu_synchronize", "foort_gpu_autotune", "foort_gpu_configure",
    "foort_gpu_queue_open", "foort_gpu_queue_close", "foort_gpu_queue_last_error", "foort_gpu_queue_add",
    "foort_gpu_queue_remove", "foort_gpu_queue_push", "foort_gpu_queue_run", "foort_gpu_queue_progress",
    "foort_gpu_queue_results", "foort_gpu_trace_positions", "foort_positions_free",
    "foort_gpu_check_division", "foort_gpu_trace_device_noalloc",
]


class FoortError(RuntimeError):
    pass


def lib():
    """Load libfoort_gpu.so (built by foort_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FoortError(f"{LIB_PATH} is not built: run python -c 'import __graft_entry__ as g; g.build()'")
        L = C.CDLL(LIB_PATH)
        L.foort_desc_size.restype = C.c_int
        L.foort_gpu_last_error.restype = C.c_char_p
        L.foort_gpu_last_error.argtypes = [C.c_void_p]
        L.foort_gpu_open.argtypes = [C.POINTER(RunDesc), C.POINTER(C.c_void_p)]
        L.foort_gpu_close.argtypes = [C.c_void_p]
        L.foort_gpu_trace_grid.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                           C.POINTER(GeodesicResults), C.POINTER(Stats)]
        L.foort_gpu_trace_pixels.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                             C.POINTER(GeodesicResults), C.POINTER(Stats)]
        L.foort_gpu_trace_device.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p,
                                             C.POINTER(GeodesicResults), C.c_void_p, C.POINTER(Stats)]
        L.foort_gpu_trace_device_noalloc.argtypes = L.foort_gpu_trace_device.argtypes
        L.foort_gpu_camera.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.foort_gpu_check_division.argtypes = [C.c_uint64] + [C.c_void_p] * 4
        L.foort_gpu_configure.argtypes = [C.c_void_p, C.POINTER(RunDesc)]
        L.foort_gpu_autotune.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint32,
                                         C.POINTER(RunDesc), C.POINTER(Stats)]
        L.foort_gpu_device_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
        L.foort_gpu_device_free.argtypes = [C.c_void_p]
        L.foort_gpu_copy_to_device.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.foort_gpu_copy_to_host.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.foort_gpu_queue_open.argtypes = [C.POINTER(RunDesc), C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]
        L.foort_gpu_queue_close.argtypes = [C.c_void_p]
        L.foort_gpu_queue_last_error.restype = C.c_char_p
        L.foort_gpu_queue_last_error.argtypes = [C.c_void_p]
        L.foort_gpu_queue_add.argtypes = [C.c_void_p, C.POINTER(RunDesc), C.c_uint64, C.POINTER(C.c_uint32)]
        L.foort_gpu_queue_remove.argtypes = [C.c_void_p, C.c_uint32]
        L.foort_gpu_queue_push.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]
        L.foort_gpu_queue_run.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(Stats)]
        L.foort_gpu_queue_progress.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64)]
        L.foort_gpu_queue_results.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(GeodesicResults)]
        L.foort_gpu_trace_positions.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                                C.POINTER(GeodesicResults), C.POINTER(Positions), C.POINTER(Stats)]
        L.foort_positions_free.argtypes = [C.POINTER(Positions)]
        if L.foort_desc_size() != C.sizeof(RunDesc):
            raise FoortError(f"foort_run_desc layout mismatch: library {L.foort_desc_size()} bytes, "
                             f"binding {C.sizeof(RunDesc)}")
        _lib = L
    return _lib


def _host_results(count):
    return {
        "term": np.zeros(count, np.int32), "quadrant": np.zeros(count, np.int32),
        "equat_passes": np.zeros(count, np.int32), "closest_radius": np.zeros(count, np.float64),
        "final_pos": np.zeros((count, 4), np.float64), "final_vel": np.zeros((count, 4), np.float64),
        "lambda": np.zeros(count, np.float64), "steps": np.zeros(count, np.uint32),
    }


def _results_struct(arrs, ptr):
    return GeodesicResults(*[ptr(arrs[k]) if arrs.get(k) is not None else None for k in
                             ("term", "quadrant", "equat_passes", "closest_radius", "final_pos",
                              "final_vel", "lambda", "steps")])


class Context:
    """One configured run (foort_gpu_open): the objects FOORT/Main.cpp builds
    from its configuration, ready to integrate batches of pixels."""

    def __init__(self, desc):
        L = lib()
        self._L = L
        self.desc = desc
        h = C.c_void_p()
        rc = L.foort_gpu_open(C.byref(desc), C.byref(h))
        if rc != 0:
            raise FoortError(f"foort_gpu_open failed ({rc}): {L.foort_gpu_last_error(None).decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.foort_gpu_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise FoortError(f"{what} failed ({rc}): {self._L.foort_gpu_last_error(self._h).decode()}")

    def trace_grid(self, side, first=0, count=None):
        """SimpleSquareMesh pixels first..first+count-1 of a side x side grid (host arrays)."""
        if count is None:
            count = side * side - first
        out = _host_results(count)
        st = Stats()
        res = _results_struct(out, lambda a: a.ctypes.data_as(C.c_void_p))
        self._check(self._L.foort_gpu_trace_grid(self._h, side, first, count, C.byref(res), C.byref(st)),
                    "foort_gpu_trace_grid")
        out["stats"] = st
        return out

    def trace_pixels(self, side, screen_index):
        """Explicit (row, column) list, as the subdivision meshes queue pixels."""
        si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
        count = si.shape[0]
        out = _host_results(count)
        st = Stats()
        res = _results_struct(out, lambda a: a.ctypes.data_as(C.c_void_p))
        self._check(self._L.foort_gpu_trace_pixels(self._h, side, count, si.ctypes.data_as(C.c_void_p),
                                                   C.byref(res), C.byref(st)), "foort_gpu_trace_pixels")
        out["stats"] = st
        return out

    def trace_device(self, side, first, count, results, screen_index=None, stream=None, stats=True,
                     noalloc=False):
        """Device-resident batch: `results` maps field name -> DeviceArray (or None);
        `screen_index` is a DeviceArray of (row, column) pairs or None for grid indices.
        noalloc: through foort_gpu_trace_device_noalloc (work memory in the library's static
        device array; the call allocates nothing)."""
        res = _results_struct(results, lambda t: C.c_void_p(t.ptr))
        si = C.c_void_p(screen_index.ptr) if screen_index is not None else None
        st = Stats() if stats else None
        fn = self._L.foort_gpu_trace_device_noalloc if noalloc else self._L.foort_gpu_trace_device
        self._check(fn(self._h, side, first, count, si, C.byref(res), C.c_void_p(stream) if stream else None,
                       C.byref(st) if st is not None else None),
                    "foort_gpu_trace_device_noalloc" if noalloc else "foort_gpu_trace_device")
        return st

    def trace_positions(self, side, screen_index=None, count=None):
        """GeodesicPosition batch (foort_gpu_trace_positions): the results of trace_pixels plus
        "offsets" (count + 1) and "points" (total x 4) of the kept positions."""
        if screen_index is not None:
            si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
            n, ptr = si.shape[0], si.ctypes.data_as(C.c_void_p)
        else:
            n, ptr = count, None
        out = _host_results(n)
        st, pos = Stats(), Positions()
        res = _results_struct(out, lambda a: a.ctypes.data_as(C.c_void_p))
        self._check(self._L.foort_gpu_trace_positions(self._h, side, n, ptr, C.byref(res), C.byref(pos), C.byref(st)),
                    "foort_gpu_trace_positions")
        try:
            out["offsets"] = np.ctypeslib.as_array(pos.offsets, shape=(n + 1,)).copy()
            out["points"] = (np.ctypeslib.as_array(pos.points, shape=(pos.total, 4)).copy() if pos.total
                             else np.zeros((0, 4)))
        finally:
            self._L.foort_positions_free(C.byref(pos))
        out["stats"] = st
        return out

    def configure(self, desc):
        """foort_gpu_configure: another run description on this context's buffers."""
        self._check(self._L.foort_gpu_configure(self._h, C.byref(desc)), "foort_gpu_configure")

    def autotune(self, side, screen_index=None, count=None, tune_schedule=False):
        """foort_gpu_autotune on a sample batch (a (row, column) list, or grid indices
        0 .. count-1); returns (the context's descriptor afterwards, the sample's Stats)."""
        chosen, st = RunDesc(), Stats()
        if screen_index is not None:
            si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
            n, ptr = si.shape[0], si.ctypes.data_as(C.c_void_p)
        else:
            n, ptr = count, None
        self._check(self._L.foort_gpu_autotune(self._h, side, n, ptr, 1 if tune_schedule else 0, C.byref(chosen),
                                               C.byref(st)), "foort_gpu_autotune")
        return chosen, st

    def camera(self):
        V, g0 = np.zeros(16), np.zeros(4)
        self._check(self._L.foort_gpu_camera(self._h, V.ctypes.data_as(C.c_void_p), g0.ctypes.data_as(C.c_void_p)),
                    "foort_gpu_camera")
        return V.reshape(4, 4), g0


class Queue:
    """Request queue (foort_gpu_queue_*): many requests of one group -- run descriptions
    that differ only in metric parameters, camera and screen -- share one device ring of
    pixels and one persistent kernel."""

    def __init__(self, group_desc, capacity=1 << 20, max_requests=256):
        L = lib()
        self._L = L
        self.max_requests = max_requests
        h = C.c_void_p()
        rc = L.foort_gpu_queue_open(C.byref(group_desc), capacity, max_requests, C.byref(h))
        if rc != 0:
            raise FoortError(f"foort_gpu_queue_open failed ({rc}): {L.foort_gpu_queue_last_error(None).decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.foort_gpu_queue_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise FoortError(f"{what} failed ({rc}): {self._L.foort_gpu_queue_last_error(self._h).decode()}")

    def add(self, desc, side):
        slot = C.c_uint32()
        self._check(self._L.foort_gpu_queue_add(self._h, C.byref(desc), side, C.byref(slot)), "foort_gpu_queue_add")
        return slot.value

    def remove(self, slot):
        self._check(self._L.foort_gpu_queue_remove(self._h, slot), "foort_gpu_queue_remove")

    def push(self, slot, screen_index):
        si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
        first = C.c_uint64()
        self._check(self._L.foort_gpu_queue_push(self._h, slot, si.shape[0], si.ctypes.data_as(C.c_void_p),
                                                 C.byref(first)), "foort_gpu_queue_push")
        return first.value, si.shape[0]

    def run(self, launch_us=0):
        st = Stats()
        self._check(self._L.foort_gpu_queue_run(self._h, launch_us, C.byref(st)), "foort_gpu_queue_run")
        return st

    def progress(self):
        """(per-slot [done, steps] array, pixels pushed, pixels done)."""
        ds = np.zeros((self.max_requests, 2), np.uint64)
        pushed, fin = C.c_uint64(), C.c_uint64()
        self._check(self._L.foort_gpu_queue_progress(self._h, ds.ctypes.data_as(C.c_void_p), C.byref(pushed),
                                                     C.byref(fin)), "foort_gpu_queue_progress")
        return ds, pushed.value, fin.value

    def results(self, first, n):
        out = {"term": np.zeros(n, np.int32), "quadrant": np.zeros(n, np.int32),
               "equat_passes": np.zeros(n, np.int32), "closest_radius": np.zeros(n, np.float64)}
        res = _results_struct(out, lambda a: a.ctypes.data_as(C.c_void_p))
        self._check(self._L.foort_gpu_queue_results(self._h, first, n, C.byref(res)), "foort_gpu_queue_results")
        return out


class DeviceArray:
    """A device allocation made through the C ABI (foort_gpu_device_alloc)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        L = lib()
        p = C.c_void_p()
        if L.foort_gpu_device_alloc(self.nbytes, C.byref(p)) != 0:
            raise FoortError(L.foort_gpu_last_error(None).decode())
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if lib().foort_gpu_copy_to_device(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), d.nbytes) != 0:
            raise FoortError(lib().foort_gpu_last_error(None).decode())
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if lib().foort_gpu_copy_to_host(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes) != 0:
            raise FoortError(lib().foort_gpu_last_error(None).decode())
        return out

    def free(self):
        if self.ptr:
            lib().foort_gpu_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


RESULT_LAYOUT = {"term": ((), np.int32), "quadrant": ((), np.int32), "equat_passes": ((), np.int32),
                 "closest_radius": ((), np.float64), "final_pos": ((4,), np.float64),
                 "final_vel": ((4,), np.float64), "lambda": ((), np.float64), "steps": ((), np.uint32)}


def device_results(count, fields=tuple(RESULT_LAYOUT)):
    """Allocate device result arrays for Context.trace_device."""
    return {k: DeviceArray((count,) + RESULT_LAYOUT[k][0], RESULT_LAYOUT[k][1]) for k in fields}


def synchronize():
    if lib().foort_gpu_synchronize() != 0:
        raise FoortError(lib().foort_gpu_last_error(None).decode())
