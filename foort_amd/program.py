"""ctypes binding of include/foort_host.h (libfoort_host.so): the host side of
the drop-in -- configuration files, meshes, output files and the program loop
of FOORT/Main.cpp around the GPU batch call.

    from foort_amd import program
    program.run("my.cfg")                       # what `foort my.cfg` does
    cfg = program.Config(text=open("my.cfg").read())
    cfg.run_desc(), cfg.mesh_desc(), cfg.first_line(), cfg.diag_names()
    mesh = program.Mesh(cfg.mesh_desc())        # Mesh::getNewInitConds / GeodesicFinished / EndCurrentLoop
"""
import ctypes as C
import os

import numpy as np

from .abi import GeodesicResults, RunDesc, Stats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FOORT_HOST_LIB") or os.path.join(_HERE, "libfoort_host.so")
CLI_PATH = os.path.join(_HERE, "foort")
_lib = None

MESH_SIMPLESQUARE = 0
MESH_INPUTCERTAINPIXELS = 1
MESH_SQUARESUBDIVISION = 2
MESH_SQUARESUBDIVISIONV2 = 3

EXPORTS = [
    "foort_host_last_error", "foort_mesh_create", "foort_mesh_destroy", "foort_mesh_count",
    "foort_mesh_row_column_size", "foort_mesh_screen_indices", "foort_mesh_finished", "foort_mesh_finished_pairs", "foort_mesh_end_loop",
    "foort_mesh_is_finished", "foort_mesh_description", "foort_config_read", "foort_config_free",
    "foort_config_run_desc", "foort_config_mesh_desc", "foort_config_first_line", "foort_config_diag_names",
    "foort_output_create", "foort_output_destroy", "foort_output_prepare", "foort_output_geodesic",
    "foort_output_finished", "foort_program_options_init", "foort_program_run",
    "foort_server_config_init", "foort_server_create", "foort_server_submit", "foort_server_results",
    "foort_server_drain", "foort_server_destroy",
]


class MeshDesc(C.Structure):
    """foort_mesh_desc (include/foort_host.h)"""
    _fields_ = [
        ("type", C.c_int32), ("value_diag", C.c_uint32), ("total_pixels", C.c_uint64),
        ("initial_pixels", C.c_uint64), ("max_pixels", C.c_uint64), ("iteration_pixels", C.c_uint64),
        ("max_subdivide", C.c_int32), ("initial_subdivision_to_final", C.c_int32),
        ("pixels", C.c_void_p), ("pixel_count", C.c_uint64),
    ]


TRACE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(RunDesc), C.c_uint64, C.c_uint64,
                       C.POINTER(C.c_uint64), C.POINTER(GeodesicResults))
POSITIONS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(RunDesc), C.c_uint64, C.c_uint64,
                           C.POINTER(C.c_uint64), C.POINTER(GeodesicResults), C.POINTER(C.c_uint64),
                           C.POINTER(C.c_double), C.c_uint64)


class ProgramOptions(C.Structure):
    """foort_program_options (include/foort_host.h)"""
    _fields_ = [
        ("output_prefix", C.c_char_p), ("christoffel", C.c_int32), ("precision", C.c_int32),
        ("schedule", C.c_int32), ("screen_output_level", C.c_int32), ("tracer", TRACE_FN),
        ("tracer_user", C.c_void_p), ("stats", C.POINTER(Stats)), ("positions_tracer", POSITIONS_FN),
    ]


class Request(C.Structure):
    """foort_request (include/foort_host.h)"""
    _fields_ = [("desc", RunDesc), ("mesh", MeshDesc), ("tag", C.c_uint64), ("output_prefix", C.c_char_p),
                ("output_extension", C.c_char_p)]


class RequestResult(C.Structure):
    """foort_request_result (include/foort_host.h)"""
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("iterations", C.c_uint32), ("pixels", C.c_uint64),
                ("steps", C.c_uint64), ("submit_s", C.c_double), ("start_s", C.c_double), ("end_s", C.c_double),
                ("kernel_ms", C.c_double), ("footprint_bytes", C.c_uint64)]


class ServerConfig(C.Structure):
    """foort_server_config (include/foort_host.h)"""
    _fields_ = [("workers", C.c_uint32), ("launch_us", C.c_uint32), ("hbm_budget_bytes", C.c_uint64)]


class HostError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HostError(f"{LIB_PATH} is not built: run python -c 'import __graft_entry__ as g; g.build()'")
        L = C.CDLL(LIB_PATH)
        L.foort_host_last_error.restype = C.c_char_p
        L.foort_mesh_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]
        L.foort_mesh_destroy.argtypes = [C.c_void_p]
        for f in ("foort_mesh_count", "foort_mesh_row_column_size"):
            getattr(L, f).restype = C.c_uint64
            getattr(L, f).argtypes = [C.c_void_p]
        L.foort_mesh_screen_indices.argtypes = [C.c_void_p, C.c_void_p]
        L.foort_mesh_finished.argtypes = [C.c_void_p, C.c_void_p]
        L.foort_mesh_finished_pairs.argtypes = [C.c_void_p, C.c_void_p]
        L.foort_mesh_end_loop.argtypes = [C.c_void_p]
        L.foort_mesh_is_finished.argtypes = [C.c_void_p]
        L.foort_mesh_description.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.foort_config_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.foort_config_free.argtypes = [C.c_void_p]
        L.foort_config_run_desc.argtypes = [C.c_void_p, C.POINTER(RunDesc)]
        L.foort_config_mesh_desc.argtypes = [C.c_void_p, C.POINTER(MeshDesc)]
        L.foort_config_first_line.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.foort_config_diag_names.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.foort_output_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p), C.c_int,
                                          C.c_uint64, C.c_uint64, C.c_char_p, C.POINTER(C.c_void_p)]
        L.foort_output_destroy.argtypes = [C.c_void_p]
        L.foort_output_prepare.argtypes = [C.c_void_p, C.c_uint64]
        L.foort_output_geodesic.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_char_p)]
        L.foort_output_finished.argtypes = [C.c_void_p]
        L.foort_program_options_init.argtypes = [C.POINTER(ProgramOptions)]
        L.foort_server_config_init.argtypes = [C.POINTER(ServerConfig)]
        L.foort_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]
        L.foort_server_submit.argtypes = [C.c_void_p, C.POINTER(Request)]
        L.foort_server_results.argtypes = [C.c_void_p, C.POINTER(RequestResult), C.c_int, C.c_double]
        L.foort_server_drain.argtypes = [C.c_void_p]
        L.foort_server_destroy.argtypes = [C.c_void_p]
        L.foort_program_run.argtypes = [C.c_char_p, C.POINTER(ProgramOptions)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HostError(lib().foort_host_last_error().decode())


def _text(fn, *args):
    n = fn(*args, None, 0)
    if n < 0:
        _check(n)
    buf = C.create_string_buffer(n + 1)
    fn(*args, buf, n + 1)
    return buf.value.decode()


class Mesh:
    """A mesh (foort_mesh_create): one of the reference's four Mesh classes."""

    def __init__(self, desc):
        self.desc = desc
        self._h = C.c_void_p()
        _check(lib().foort_mesh_create(C.byref(desc), C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().foort_mesh_destroy(self._h)
            self._h = None

    @property
    def count(self):
        return lib().foort_mesh_count(self._h)

    @property
    def side(self):
        return lib().foort_mesh_row_column_size(self._h)

    def screen_indices(self):
        out = np.zeros((self.count, 2), np.uint64)
        _check(lib().foort_mesh_screen_indices(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def finished(self, values):
        v = np.ascontiguousarray(values, np.float64)
        assert v.shape == (self.count,)
        _check(lib().foort_mesh_finished(self._h, v.ctypes.data_as(C.c_void_p)))

    def finished_pairs(self, pairs):
        """GeodesicPosition mesh values: (theta, phi) per geodesic."""
        v = np.ascontiguousarray(pairs, np.float64)
        assert v.shape == (self.count, 2)
        _check(lib().foort_mesh_finished_pairs(self._h, v.ctypes.data_as(C.c_void_p)))

    def end_loop(self):
        _check(lib().foort_mesh_end_loop(self._h))

    @property
    def is_finished(self):
        return bool(lib().foort_mesh_is_finished(self._h))

    def description(self):
        return _text(lib().foort_mesh_description, self._h)


def mesh_desc(type, value_diag, total_pixels=0, initial_pixels=0, max_pixels=0, iteration_pixels=0,
              max_subdivide=1, initial_subdivision_to_final=False, pixels=None):
    """A foort_mesh_desc; `pixels` ((row, column) pairs) is InputCertainPixelsMesh's list, kept
    alive with the descriptor."""
    m = MeshDesc(type=type, value_diag=value_diag, total_pixels=total_pixels, initial_pixels=initial_pixels,
                 max_pixels=max_pixels, iteration_pixels=iteration_pixels, max_subdivide=max_subdivide,
                 initial_subdivision_to_final=int(initial_subdivision_to_final))
    if pixels is not None:
        px = np.ascontiguousarray(pixels, dtype=np.uint64).reshape(-1, 2)
        m._pixels = px
        m.pixels = px.ctypes.data
        m.pixel_count = px.shape[0]
    return m


class Config:
    """A configuration file read the way FOORT/Config.cpp reads it (foort_config_read)."""

    def __init__(self, path=None, text=None):
        self._h = C.c_void_p()
        _check(lib().foort_config_read(path.encode() if path else None, text.encode() if text is not None else None,
                                       C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().foort_config_free(self._h)
            self._h = None

    def run_desc(self):
        d = RunDesc()
        _check(lib().foort_config_run_desc(self._h, C.byref(d)))
        return d

    def mesh_desc(self):
        m = MeshDesc()
        _check(lib().foort_config_mesh_desc(self._h, C.byref(m)))
        return m

    def first_line(self):
        return _text(lib().foort_config_first_line, self._h)

    def diag_names(self):
        return _text(lib().foort_config_diag_names, self._h).split()


class Output:
    """GeodesicOutputHandler (foort_output_create)."""

    def __init__(self, prefix, names, timestamp="", extension="", to_cache=2 ** 64 - 2, per_file=0, first_line=""):
        self._names = (C.c_char_p * len(names))(*[n.encode() for n in names])
        self._h = C.c_void_p()
        _check(lib().foort_output_create(prefix.encode(), timestamp.encode(), extension.encode(), self._names,
                                         len(names), to_cache, per_file, first_line.encode(), C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().foort_output_destroy(self._h)
            self._h = None

    def prepare(self, n):
        _check(lib().foort_output_prepare(self._h, n))

    def geodesic(self, index, row, col, texts):
        arr = (C.c_char_p * (len(texts) + 1))(*[t.encode() for t in texts], None)
        _check(lib().foort_output_geodesic(self._h, index, row, col, arr))

    def finished(self):
        _check(lib().foort_output_finished(self._h))


_METRIC_CFG = {0: ("Kerr", ("a",)), 1: ("FlatSpace", ()), 2: ("RasheedLarsen", ("m", "a", "p", "q")),
               3: ("Johannsen", ("a", "alpha13", "alpha22", "alpha52", "epsilon3")),
               4: ("MankoNovikov", ("a", "alpha3"))}
_MESH_CFG = {MESH_SIMPLESQUARE: "SimpleSquareMesh", MESH_INPUTCERTAINPIXELS: "InputCertainPixelsMesh",
             MESH_SQUARESUBDIVISION: "SquareSubdivisionMesh", MESH_SQUARESUBDIVISIONV2: "SquareSubdivisionMeshV2"}


def config_text(desc, mesh, prefix=None, extension="dat"):
    """A configuration file (the reference's option names, FOORT/Config.cpp) that the program
    reads back into this run description and mesh; output files `prefix`_<name>.`extension`
    without a time stamp (console output when prefix is None)."""
    from . import abi as A

    def num(x):
        return repr(float(x))

    def cnt(n):
        n = int(n)
        return f"{n}L" if n >= 2 ** 31 else str(n)

    def upd(u, gp=False):
        return (f"UpdateFrequency = {cnt(u.update_n_steps)}; UpdateStart = {'true' if u.update_start else 'false'}; "
                f"UpdateFinish = {'true' if u.update_finish else 'false'};")

    name, pnames = _METRIC_CFG[desc.metric]
    params = "".join(f" {n} = {num(desc.metric_params[i])};" for i, n in enumerate(pnames))
    diag = []
    for flag, dname, extra in ((A.DIAGFLAG_FOURCOLORSCREEN, "FourColorScreen", ""),
                               (A.DIAGFLAG_GEODESICPOSITION, "GeodesicPosition",
                                f" {upd(desc.gp_update)} OutputSteps = {cnt(desc.gp_output_steps)};"),
                               (A.DIAGFLAG_EQUATORIALPASSES, "EquatorialPasses",
                                f" {upd(desc.ep_update)} Threshold = {num(desc.ep_threshold)};"),
                               (A.DIAGFLAG_CLOSESTRADIUS, "ClosestRadius", f" {upd(desc.cr_update)}")):
        if desc.diag_flags & flag:
            use = " UseForMesh = true;" if desc.value_diag == flag else ""
            diag.append(f"  {dname} = {{ On = true;{use}{extra} }};")
    term = []
    tf = desc.term_flags
    if tf & A.TERMFLAG_HORIZON:
        term.append(f"  Horizon = {{ On = true; Epsilon_Horizon = {num(desc.horizon_epsilon)}; "
                    f"UpdateFrequency = {cnt(desc.horizon_update_n)}; }};")
    if tf & A.TERMFLAG_BOUNDARYSPHERE:
        term.append(f"  BoundarySphere = {{ On = true; SphereRadius = {num(desc.sphere_radius)}; "
                    f"UpdateFrequency = {cnt(desc.sphere_update_n)}; }};")
    if tf & A.TERMFLAG_TIMEOUT:
        term.append(f"  TimeOut = {{ On = true; MaxSteps = {cnt(desc.timeout_max_steps)}; "
                    f"UpdateFrequency = {cnt(desc.timeout_update_n)}; }};")
    if tf & A.TERMFLAG_THETASINGULARITY:
        term.append(f"  ThetaSingularity = {{ On = true; Epsilon = {num(desc.theta_epsilon)}; "
                    f"UpdateFrequency = {cnt(desc.theta_update_n)}; }};")
    mtype = _MESH_CFG[mesh.type]
    if mesh.type in (MESH_SIMPLESQUARE, MESH_INPUTCERTAINPIXELS):
        mopts = f"TotalPixels = {cnt(mesh.total_pixels)};"
    else:
        mopts = (f"InitialPixels = {cnt(mesh.initial_pixels)}; MaxPixels = {cnt(mesh.max_pixels)}; "
                 f"IterationPixels = {cnt(mesh.iteration_pixels)}; MaxSubdivide = {int(mesh.max_subdivide)}; "
                 f"InitialSubdivisionToFinal = {'true' if mesh.initial_subdivision_to_final else 'false'};")
    cam = desc.camera_position
    integ = "Verlet" if desc.integrator == A.INTEGRATOR_VERLET else "RK4"
    chris = "Analytic" if desc.christoffel == A.CHRISTOFFEL_ANALYTIC else "FD"
    prec = "Mixed" if desc.precision == A.PRECISION_MIXED else "FP64"
    out = (f'FilePrefix = "{prefix}"; FileExtension = "{extension}"; TimeStamp = false; FirstLineInfo = true;'
           if prefix else "")
    return f"""Metric = {{ Name = "{name}";{params} RLogScale = {'true' if desc.rlog_scale else 'false'}; }};
Source = {{ Name = "NoSource"; }};
Diagnostics = {{
{chr(10).join(diag)}
}};
Terminations = {{
{chr(10).join(term)}
}};
ViewScreen = {{
  Position = {{ t = {num(cam[0])}; r = {num(cam[1])}; theta = {num(cam[2])}; phi = {num(cam[3])}; }};
  ScreenSize = {{ x = {num(desc.screen_size[0])}; y = {num(desc.screen_size[1])}; }};
  ScreenCenter = {{ x = {num(desc.screen_center[0])}; y = {num(desc.screen_center[1])}; }};
  Mesh = {{ Type = "{mtype}"; {mopts} }};
}};
Integrator = {{ Type = "{integ}"; StepSize = {num(desc.step_size)}; DerivativeH = {num(desc.derivative_h)};
  SmallestPossibleStepsize = {num(desc.smallest_step)}; VerletVelocityTolerance = {num(desc.verlet_tolerance)}; }};
GPU = {{ Christoffel = "{chris}"; Precision = "{prec}"; StepCap = {cnt(desc.step_cap)};
  MixedSwitchRadius = {num(desc.mixed_switch_radius)}; MixedPoleMargin = {num(desc.mixed_pole_margin)}; }};
Output = {{ ScreenOutputLevel = 1; {out} }};
"""


def run(cfg_path, output_prefix=None, christoffel=-1, precision=-1, schedule=-1, screen_output_level=-1,
        tracer=None):
    """Run the program (foort_program_run) on a configuration file; returns the GPU totals
    (foort_stats).  `tracer(desc, side, screen_index[n, 2]) -> dict of numpy arrays` replaces the
    GPU batch call when given (the tests drive the driver with the CPU oracle this way)."""
    o = ProgramOptions()
    lib().foort_program_options_init(C.byref(o))
    o.output_prefix = output_prefix.encode() if output_prefix is not None else None
    o.christoffel, o.precision, o.schedule = christoffel, precision, schedule
    o.screen_output_level = screen_output_level
    st = Stats()
    o.stats = C.pointer(st)
    if tracer is not None:
        def _fill(out_p, count, res):
            out = out_p.contents
            for name, ctype in (("term", C.c_int32), ("quadrant", C.c_int32), ("equat_passes", C.c_int32),
                                ("closest_radius", C.c_double)):
                ptr = getattr(out, name)
                if ptr:
                    dst = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,))
                    dst[:] = res[name]

        def _cb(user, desc_p, side, count, si_p, out_p):
            try:
                si = np.ctypeslib.as_array(si_p, shape=(count * 2,)).reshape(count, 2).copy()
                _fill(out_p, count, tracer(desc_p.contents, side, si))
                return 0
            except Exception:  # noqa: BLE001 -- reported through the return code
                import traceback
                traceback.print_exc()
                return -1

        def _pos_cb(user, desc_p, side, count, si_p, out_p, off_p, pts_p, cap):
            # the tracer returns "offsets" and "points" too when the description has GeodesicPosition
            try:
                si = np.ctypeslib.as_array(si_p, shape=(count * 2,)).reshape(count, 2).copy()
                res = tracer(desc_p.contents, side, si)
                _fill(out_p, count, res)
                np.ctypeslib.as_array(off_p, shape=(count + 1,))[:] = res["offsets"]
                total = int(res["offsets"][-1])
                if total and total <= cap:
                    np.ctypeslib.as_array(pts_p, shape=(total * 4,))[:] = np.asarray(res["points"]).reshape(-1)
                return 0
            except Exception:  # noqa: BLE001 -- reported through the return code
                import traceback
                traceback.print_exc()
                return -1
        o.tracer = TRACE_FN(_cb)
        o.positions_tracer = POSITIONS_FN(_pos_cb)
    rc = lib().foort_program_run(cfg_path.encode(), C.byref(o))
    _check(rc)
    return st


class Server:
    """The render server (foort_server_*): submit requests, collect their results."""

    def __init__(self, workers=0, hbm_budget_bytes=0, launch_us=0):
        cfg = ServerConfig()
        lib().foort_server_config_init(C.byref(cfg))
        if workers:
            cfg.workers = workers
        if launch_us:
            cfg.launch_us = launch_us
        if hbm_budget_bytes:
            cfg.hbm_budget_bytes = hbm_budget_bytes
        self.config = cfg
        self._h = C.c_void_p()
        _check(lib().foort_server_create(C.byref(cfg), C.byref(self._h)))

    def submit(self, desc, mesh, tag, output_prefix=None, output_extension=None):
        """Queue a request; with output_prefix it writes the reference's output files."""
        req = Request(desc=desc, mesh=mesh, tag=tag,
                      output_prefix=output_prefix.encode() if output_prefix else None,
                      output_extension=output_extension.encode() if output_extension else None)
        _check(lib().foort_server_submit(self._h, C.byref(req)))

    def results(self, cap=4096, timeout=0.0):
        arr = (RequestResult * cap)()
        n = lib().foort_server_results(self._h, arr, cap, timeout)
        if n < 0:
            _check(n)
        return [arr[i] for i in range(n)]

    def drain(self):
        _check(lib().foort_server_drain(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().foort_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()
