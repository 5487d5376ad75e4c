"""TEST INFRASTRUCTURE ONLY: ctypes loaders for the CPU oracle
(liboracle_libm.so, liboracle_fm.so) and for the reference's own path built
into oracle/_ref.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference results to compare with -- never by
the product.
"""
import ctypes as C
import os

import numpy as np

from foort_amd.abi import RunDesc

HERE = os.path.dirname(os.path.abspath(__file__))

_LIBS = {}


def _load(name):
    if name not in _LIBS:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        lib = C.CDLL(path)
        _LIBS[name] = lib
    return _LIBS[name]


def lib(kind):
    """kind: 'libm' (C library math, pinned to the reference), 'fm' (portable math,
    what the GPU path is compared with bit for bit), 'ref' (the reference sources, IEEE flags), 'ref_ofast'
    (the reference sources with their own -Ofast flags)."""
    return _load({"libm": "liboracle_libm.so", "fm": "liboracle_fm.so",
                  "ref": "_ref/libfoort_ref.so", "ref_ofast": "_ref/libfoort_ref_ofast.so"}[kind])


def have(kind):
    try:
        lib(kind)
        return True
    except OSError:
        return False


def trace(kind, desc, size, first=0, count=None, screen_index=None, nthreads=None):
    """Integrate pixels with an oracle variant; returns a dict of numpy arrays."""
    L = lib(kind)
    fn = L.oracle_trace if kind in ("libm", "fm") else L.ref_trace
    if screen_index is not None:
        si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
        count = si.shape[0]
        si_ptr = si.ctypes.data_as(C.c_void_p)
    else:
        if count is None:
            count = size * size - first
        si_ptr = None
    out = {
        "term": np.zeros(count, np.int32), "quadrant": np.zeros(count, np.int32),
        "equat_passes": np.zeros(count, np.int32), "closest_radius": np.zeros(count, np.float64),
        "final_pos": np.zeros((count, 4), np.float64), "final_vel": np.zeros((count, 4), np.float64),
        "lambda": np.zeros(count, np.float64), "steps": np.zeros(count, np.uint64),
    }
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    fn.argtypes = [C.POINTER(RunDesc), C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p] + [C.c_void_p] * 8 + [C.c_int]
    fn(C.byref(desc), size, first, count, si_ptr,
       *[out[k].ctypes.data_as(C.c_void_p) for k in
         ("term", "quadrant", "equat_passes", "closest_radius", "final_pos", "final_vel", "lambda", "steps")],
       nthreads)
    return out


def trace_positions(kind, desc, size, first=0, count=None, screen_index=None, nthreads=None):
    """oracle_trace with GeodesicPosition ("libm" or "fm"): the results of trace() plus
    "kept" (points per pixel), "offsets" (count + 1) and "points" (total x 4)."""
    L = lib(kind)
    fn = L.oracle_trace_positions
    if screen_index is not None:
        si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
        count = si.shape[0]
        si_ptr = si.ctypes.data_as(C.c_void_p)
    else:
        if count is None:
            count = size * size - first
        si_ptr = None
    out = {
        "term": np.zeros(count, np.int32), "quadrant": np.zeros(count, np.int32),
        "equat_passes": np.zeros(count, np.int32), "closest_radius": np.zeros(count, np.float64),
        "final_pos": np.zeros((count, 4), np.float64), "final_vel": np.zeros((count, 4), np.float64),
        "lambda": np.zeros(count, np.float64), "steps": np.zeros(count, np.uint64),
        "kept": np.zeros(count, np.uint64),
    }
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    fn.restype = C.c_uint64
    fn.argtypes = ([C.POINTER(RunDesc), C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p] + [C.c_void_p] * 8 +
                   [C.c_int, C.c_void_p, C.c_void_p, C.c_uint64])
    args = [C.byref(desc), size, first, count, si_ptr] + [out[k].ctypes.data_as(C.c_void_p) for k in
            ("term", "quadrant", "equat_passes", "closest_radius", "final_pos", "final_vel", "lambda", "steps")]
    total = fn(*args, nthreads, out["kept"].ctypes.data_as(C.c_void_p), None, 0)
    points = np.zeros((max(total, 1), 4), np.float64)
    total2 = fn(*args, nthreads, out["kept"].ctypes.data_as(C.c_void_p), points.ctypes.data_as(C.c_void_p), total)
    assert total2 == total
    out["points"] = points[:total]
    out["offsets"] = np.concatenate([[0], np.cumsum(out["kept"])]).astype(np.uint64)
    return out


def _vec(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def metric(kind, desc, p):
    L = lib(kind)
    fn = L.oracle_metric if kind in ("libm", "fm") else L.ref_metric
    p = _vec(p)
    gdd = np.zeros(16)
    guu = np.zeros(16)
    fn(C.byref(desc), p.ctypes.data_as(C.c_void_p), gdd.ctypes.data_as(C.c_void_p), guu.ctypes.data_as(C.c_void_p))
    return gdd.reshape(4, 4), guu.reshape(4, 4)


def christoffel(kind, desc, p):
    L = lib(kind)
    fn = L.oracle_christoffel if kind in ("libm", "fm") else L.ref_christoffel
    p = _vec(p)
    G = np.zeros(64)
    fn(C.byref(desc), p.ctypes.data_as(C.c_void_p), G.ctypes.data_as(C.c_void_p))
    return G.reshape(4, 4, 4)


def step(kind, desc, x, v):
    L = lib(kind)
    fn = L.oracle_step if kind in ("libm", "fm") else L.ref_step
    x, v = _vec(x), _vec(v)
    nx, nv, h = np.zeros(4), np.zeros(4), C.c_double()
    fn(C.byref(desc), x.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
       nx.ctypes.data_as(C.c_void_p), nv.ctypes.data_as(C.c_void_p), C.byref(h))
    return nx, nv, h.value


def initial(kind, desc, size, row, col):
    L = lib(kind)
    pos, vel = np.zeros(4), np.zeros(4)
    if kind in ("libm", "fm"):
        L.oracle_initial.argtypes = [C.POINTER(RunDesc), C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.oracle_initial(C.byref(desc), row * 1.0 / float(size - 1), col * 1.0 / float(size - 1),
                         pos.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p))
    else:
        L.ref_initial.argtypes = [C.POINTER(RunDesc), C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]
        L.ref_initial(C.byref(desc), size, row, col, pos.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p))
    return pos, vel


def camera(kind, desc):
    L = lib(kind)
    V, g0 = np.zeros(16), np.zeros(4)
    L.oracle_camera(C.byref(desc), V.ctypes.data_as(C.c_void_p), g0.ctypes.data_as(C.c_void_p))
    return V.reshape(4, 4), g0


def horizon_radius(kind, desc):
    L = lib(kind)
    L.oracle_horizon_radius.restype = C.c_double
    return L.oracle_horizon_radius(C.byref(desc))


def time_reference(kind, desc, size, screen_index, nthreads):
    """Run the reference's per-pixel loop once per pixel (oracle/_ref, `kind`
    'ref' or 'ref_ofast'); returns (total steps, wall seconds)."""
    import time
    L = lib(kind)
    si = np.ascontiguousarray(screen_index, dtype=np.uint64).reshape(-1, 2)
    L.ref_time.restype = C.c_ulonglong
    L.ref_time.argtypes = [C.POINTER(RunDesc), C.c_uint64, C.c_uint64, C.c_void_p, C.c_int]
    t = time.perf_counter()
    steps = L.ref_time(C.byref(desc), size, si.shape[0], si.ctypes.data_as(C.c_void_p), nthreads)
    return int(steps), time.perf_counter() - t


def ref_mesh_run(mesh_type, pixels, max_pixels, iteration, max_sub, to_final, valdiag, value_grid):
    """Drive one of the reference's meshes (oracle/_ref) with values read from
    value_grid[row, col]; returns (screen indices [n, 2] in queue order, pixels per
    iteration, the mesh's description string)."""
    L = lib("ref")
    grid = np.ascontiguousarray(value_grid, dtype=np.float64)
    side = grid.shape[0]
    cap = side * side
    out = np.zeros((cap, 2), np.uint64)
    iters = np.zeros(4096, np.uint64)
    nit = C.c_int()
    desc = C.create_string_buffer(1024)
    L.ref_mesh_run.restype = C.c_longlong
    L.ref_mesh_run.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_uint32,
                               C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int,
                               C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
    n = L.ref_mesh_run(mesh_type, pixels, max_pixels, iteration, max_sub, int(to_final), valdiag,
                       grid.ctypes.data_as(C.c_void_p), side, out.ctypes.data_as(C.c_void_p), cap,
                       iters.ctypes.data_as(C.c_void_p), len(iters), C.byref(nit), desc, len(desc))
    if n < 0:
        raise RuntimeError(f"ref_mesh_run failed ({n})")
    return out[:n], iters[:nit.value], desc.value.decode()


def ref_first_line(desc, mesh_type, pixels, max_pixels=0, iteration=0, max_sub=1, to_final=False):
    """Utilities::GetFirstLineInfoString of the reference for a run description and mesh."""
    L = lib("ref")
    buf = C.create_string_buffer(4096)
    L.ref_first_line.argtypes = [C.POINTER(RunDesc), C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                                 C.c_int, C.c_char_p, C.c_size_t]
    L.ref_first_line(C.byref(desc), mesh_type, pixels, max_pixels, iteration, max_sub, int(to_final), buf, len(buf))
    return buf.value.decode()


def ref_output_run(prefix, names, timestamp, ext, to_cache, per_file, first_line, batches):
    """Feed the reference GeodesicOutputHandler batches of synthetic geodesic outputs."""
    L = lib("ref")
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    b = np.ascontiguousarray(batches, dtype=np.uint64)
    L.ref_output_run.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.c_uint64,
                                 C.c_uint64, C.c_char_p, C.c_void_p, C.c_int]
    L.ref_output_run(prefix.encode(), timestamp.encode(), ext.encode(), arr, len(names), to_cache, per_file,
                     first_line.encode(), b.ctypes.data_as(C.c_void_p), len(b))


def ref_program_run(desc, mesh_type, pixels, prefix, ext="dat", max_pixels=0, iteration=0, max_sub=1,
                    to_final=False, to_cache=2 ** 64 - 2, per_file=2 ** 64 - 1, first_line=True, nthreads=None):
    """The reference program loop (oracle/_ref) with one of its meshes, writing output files."""
    L = lib("ref")
    L.ref_program_run.restype = C.c_longlong
    L.ref_program_run.argtypes = [C.POINTER(RunDesc), C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                  C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int]
    return L.ref_program_run(C.byref(desc), mesh_type, pixels, max_pixels, iteration, max_sub, int(to_final),
                             prefix.encode(), ext.encode(), to_cache, per_file, int(first_line),
                             nthreads or min(8, os.cpu_count() or 1))
