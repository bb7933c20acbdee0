"""ctypes binding of the MI355X C ABI (include/fmt2d_hip.h).

The product path runs only through libfmt2d_hip.so; there is no CPU
fallback.  Loading fails loudly when the library is missing or was built
for another architecture.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import sys
from typing import Optional, Sequence

import numpy as np

from . import formats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfmt2d_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# fmt2d_forward flags (include/fmt2d_hip.h)
FWD_NO_HEADWAVE = 1
FWD_PROGRAM = 2


class GridInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("x0", C.c_double), ("z0", C.c_double), ("region_nnode", C.c_int * 7)]


class ForwardStats(C.Structure):
    _fields_ = [("nsrc", C.c_int), ("nfield", C.c_int), ("launches", C.c_int),
                ("tiles", C.c_longlong), ("region_nodes", C.c_longlong), ("ms_total", C.c_double),
                ("ms_fim", C.c_double), ("exit_nodes", C.c_longlong), ("sweeps", C.c_longlong),
                ("capped", C.c_longlong), ("max_sweeps", C.c_int), ("dispatches", C.c_int)]


class FieldSummary(C.Structure):
    _fields_ = [("reached", C.c_longlong), ("hash", C.c_ulonglong), ("max_resid", C.c_double),
                ("mean_resid", C.c_double), ("n_resid", C.c_longlong), ("n_resid_big", C.c_longlong)]


class AdjointStats(C.Structure):
    _fields_ = [("fields", C.c_int), ("rounds", C.c_int), ("leaves_unmatched", C.c_longlong),
                ("dropped_partials", C.c_longlong), ("ms_record", C.c_double), ("ms_reverse", C.c_double),
                ("ms_total", C.c_double), ("peak_bytes", C.c_longlong)]


class TuneResult(C.Structure):
    _fields_ = [("tile_x", C.c_int), ("tile_z", C.c_int), ("max_sweeps", C.c_int), ("max_sweeps_heavy", C.c_int),
                ("colors", C.c_int), ("window_tiles", C.c_double), ("ms_best", C.c_double), ("ms_start", C.c_double),
                ("trials", C.c_int)]


class SystemParams(C.Structure):
    _fields_ = [("crit_chi", C.c_double), ("phase_wt_psed", C.c_double), ("phase_wt_pmp", C.c_double),
                ("phase_wt_pgss", C.c_double), ("refl_weight", C.c_double), ("alt_mode", C.c_int)]


class LsqrParams(C.Structure):
    _fields_ = [("wdv", C.c_double), ("wdv_iface", C.c_double), ("w_smooth", C.c_double),
                ("w_smooth_iface", C.c_double), ("smooth_type", C.c_int), ("max_iter", C.c_int),
                ("corr_x", C.c_double), ("corr_z", C.c_double), ("corr_x_iface", C.c_double)]


class LsqrStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rnorm", C.c_double), ("bnorm", C.c_double), ("test2", C.c_double),
                ("ms", C.c_double)]


class DampingParams(C.Structure):
    _fields_ = [("lam_damp", C.c_double), ("lam_damp_iface", C.c_double), ("target_dv", C.c_double),
                ("target_dd", C.c_double), ("max_rel_pert", C.c_double)]


class DampingResult(C.Structure):
    _fields_ = [("wd", C.c_double), ("wd_iface", C.c_double), ("avg_dmv", C.c_double), ("avg_dmd", C.c_double),
                ("solves", C.c_int), ("iterations", C.c_int)]


class LineSearchParams(C.Structure):
    _fields_ = [("crit_chi", C.c_double), ("iface_step_ratio", C.c_double), ("first_step", C.c_int)]


class LineSearchResult(C.Structure):
    _fields_ = [("step", C.c_double), ("rms_orig", C.c_double), ("rms_new", C.c_double),
                ("next_first_step", C.c_int), ("forwards", C.c_int)]


class Fmt2dError(RuntimeError):
    pass


_LIB = None


def load() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise Fmt2dError(f"{LIB_PATH} is missing: build it with `python -m fmt2d_amd.build`")
    # One HIP runtime per process: torch's ROCm build carries its own
    # libamdhip64 under the same soname, and whichever is loaded first serves
    # both; torch does not initialise on another build's runtime, while this
    # library runs on torch's.  So torch, when installed, is loaded first.
    if "torch" not in sys.modules and importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    lib.fmt2d_last_error.restype = C.c_char_p
    lib.fmt2d_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    lib.fmt2d_ctx_destroy.argtypes = [C.c_void_p]
    lib.fmt2d_set_model.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                    C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    lib.fmt2d_grid_info_get.argtypes = [C.c_void_p, C.POINTER(GridInfo)]
    lib.fmt2d_forward.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int]
    lib.fmt2d_forward_stats_get.argtypes = [C.c_void_p, C.POINTER(ForwardStats)]
    lib.fmt2d_forward_summary.argtypes = [C.c_void_p, C.POINTER(FieldSummary), C.c_int, _ip]
    lib.fmt2d_forward_trace.argtypes = [C.c_void_p, _ip, _ip, C.c_int, _ip]
    lib.fmt2d_forward_sweep_hist.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    lib.fmt2d_set_tuning.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
    lib.fmt2d_tf_region.argtypes = [C.c_int]
    lib.fmt2d_get_field.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
    lib.fmt2d_set_inversion.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip]
    lib.fmt2d_inversion_info.argtypes = [C.c_void_p, _ip, _ip]
    lib.fmt2d_set_inversion_mesh.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.fmt2d_set_frechet_chunk.argtypes = [C.c_void_p, C.c_longlong]
    lib.fmt2d_set_heavy_sweeps.argtypes = [C.c_void_p, C.c_int]
    lib.fmt2d_set_launch_cap.argtypes = [C.c_void_p, C.c_int]
    lib.fmt2d_set_solver_launch.argtypes = [C.c_void_p, C.c_int]
    lib.fmt2d_set_tile.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.fmt2d_set_lsqr_mode.argtypes = [C.c_void_p, C.c_int]
    lib.fmt2d_ordered_sumsq.argtypes = [C.c_void_p, _dp, C.c_longlong, _dp]
    lib.fmt2d_set_field_precision.argtypes = [C.c_void_p, C.c_int]
    lib.fmt2d_field_snapshot.argtypes = [C.c_void_p]
    lib.fmt2d_field_compare.argtypes = [C.c_void_p, C.c_double, _dp, C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_longlong), C.c_int, _ip]
    lib.fmt2d_set_ray_search.argtypes = [C.c_void_p, C.c_int]
    lib.fmt2d_ray_nearest_calls.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    lib.fmt2d_ray_nearest_check.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _ip, _ip]
    lib.fmt2d_autotune.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, C.c_int, C.POINTER(TuneResult)]
    lib.fmt2d_mesh_info.argtypes = [C.c_void_p, _ip, _ip, _ip]
    lib.fmt2d_trace.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, _dp, _dp, _ip]
    lib.fmt2d_ray_info.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _ip]
    lib.fmt2d_ray_points.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_int, _ip]
    lib.fmt2d_frechet_csr.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, C.c_int, _ip]
    lib.fmt2d_build_system.argtypes = [C.c_void_p, _dp, C.POINTER(SystemParams), _ip, _ip, _dp]
    lib.fmt2d_system_get.argtypes = [C.c_void_p, _dp, _dp]
    lib.fmt2d_lsqr.argtypes = [C.c_void_p, C.POINTER(LsqrParams), _dp, _dp, C.POINTER(LsqrStats)]
    lib.fmt2d_damped_update.argtypes = [C.c_void_p, C.POINTER(LsqrParams), C.POINTER(DampingParams), _dp, _dp,
                                        C.POINTER(DampingResult)]
    lib.fmt2d_apply_model_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
    lib.fmt2d_model_get.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip]
    lib.fmt2d_model_put.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int]
    lib.fmt2d_frechet_apply.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.fmt2d_frechet_apply_noalloc.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.fmt2d_adjoint_gradient.argtypes = [C.c_void_p, _dp, _dp, _dp, C.POINTER(AdjointStats)]
    lib.fmt2d_set_adjoint_memory.argtypes = [C.c_void_p, C.c_longlong]
    lib.fmt2d_param_index.argtypes = [C.c_void_p, _ip, C.c_int, _ip]
    lib.fmt2d_line_search.argtypes = [C.c_void_p, C.POINTER(LineSearchParams), _dp, _dp, _ip,
                                      C.POINTER(LineSearchResult)]
    _LIB = lib
    return lib


# exported symbols declared in include/fmt2d_hip.h
EXPORTS = ("fmt2d_ctx_create", "fmt2d_ctx_destroy", "fmt2d_last_error", "fmt2d_set_model",
           "fmt2d_grid_info_get", "fmt2d_forward", "fmt2d_forward_stats_get", "fmt2d_forward_trace",
           "fmt2d_forward_sweep_hist", "fmt2d_forward_summary",
           "fmt2d_set_tuning", "fmt2d_set_heavy_sweeps", "fmt2d_set_launch_cap", "fmt2d_set_solver_launch",
           "fmt2d_set_tile",
           "fmt2d_autotune", "fmt2d_set_lsqr_mode", "fmt2d_ordered_sumsq", "fmt2d_set_field_precision", "fmt2d_field_snapshot",
           "fmt2d_field_compare", "fmt2d_set_ray_search", "fmt2d_ray_nearest_calls",
           "fmt2d_ray_nearest_check",
           "fmt2d_tf_region",
           "fmt2d_get_field",
           "fmt2d_set_inversion", "fmt2d_set_inversion_mesh", "fmt2d_set_frechet_chunk", "fmt2d_inversion_info", "fmt2d_mesh_info",
           "fmt2d_trace", "fmt2d_ray_info", "fmt2d_ray_points",
           "fmt2d_frechet_csr", "fmt2d_frechet_apply", "fmt2d_frechet_apply_noalloc",
           "fmt2d_adjoint_gradient", "fmt2d_set_adjoint_memory",
           "fmt2d_param_index", "fmt2d_build_system", "fmt2d_system_get", "fmt2d_lsqr",
           "fmt2d_damped_update", "fmt2d_model_get", "fmt2d_model_put", "fmt2d_apply_model_step", "fmt2d_line_search")


def _check(rc: int):
    if rc != 0:
        raise Fmt2dError(f"fmt2d error {rc}: {load().fmt2d_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(_dp)


class Context:
    """Device context: model, resident traveltime fields."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = C.c_void_p()
        _check(self._lib.fmt2d_ctx_create(device, C.byref(h)))
        self._h = h
        self.info: Optional[GridInfo] = None
        self._nray = 0

    def close(self):
        if self._h:
            self._lib.fmt2d_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, smesh: formats.Smesh, sed_x, sed_z, moho_x, moho_z):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (smesh.xpos, smesh.topo, smesh.zrel, smesh.vel, sed_x, sed_z, moho_x, moho_z)]
        xp, tp, zr, vl, sx, sz, mx, mz = keep
        _check(self._lib.fmt2d_set_model(self._h, len(xp), len(zr), _p(xp), _p(tp), _p(zr), _p(vl),
                                         float(smesh.vwater), len(sx), _p(sx), _p(sz), len(mx), _p(mx), _p(mz)))
        info = GridInfo()
        _check(self._lib.fmt2d_grid_info_get(self._h, C.byref(info)))
        self.info = info

    def forward(self, src_x: Sequence[float], headwave: bool = True, program: bool = False):
        """All fields of every source.  program: the forward program's variant
        (FMT2D_FWD_PROGRAM: its reflection normals, no head-wave pass, its ray
        rules in the next trace)."""
        xs = np.ascontiguousarray(src_x, dtype=np.float64)
        flags = (0 if headwave else FWD_NO_HEADWAVE) | (FWD_PROGRAM if program else 0)
        self._nray = 0   # a forward drops the rays of the last trace
        _check(self._lib.fmt2d_forward(self._h, len(xs), _p(xs), flags))

    def set_tuning(self, window_tiles: float = 3.0, max_sweeps: int = 16, colors: int = 2):
        _check(self._lib.fmt2d_set_tuning(self._h, float(window_tiles), int(max_sweeps), int(colors)))

    def set_heavy_sweeps(self, max_sweeps_heavy: int = 8):
        _check(self._lib.fmt2d_set_heavy_sweeps(self._h, int(max_sweeps_heavy)))

    def set_tile(self, tile_x: int = 8, tile_z: int = 8):
        _check(self._lib.fmt2d_set_tile(self._h, int(tile_x), int(tile_z)))

    def autotune(self, src_x, flags: int = 0, level: int = 0) -> TuneResult:
        """Pick the fastest tile schedule for these sources (fmt2d_autotune);
        the context keeps it and ends solved for the sources."""
        x = np.ascontiguousarray(src_x, dtype=np.float64)
        r = TuneResult()
        self._nray = 0   # a forward drops the rays of the last trace
        _check(self._lib.fmt2d_autotune(self._h, len(x), _p(x), int(flags), int(level), C.byref(r)))
        return r

    def set_field_precision(self, bits: int = 64):
        """Storage precision of the tile solver's grid-node fields (64 or 32)."""
        _check(self._lib.fmt2d_set_field_precision(self._h, int(bits)))

    def field_snapshot(self):
        _check(self._lib.fmt2d_field_snapshot(self._h))

    def field_compare(self, tol: float = 1e-4):
        """(max |dT| per field, slots reached by only one, slots with
        |dT| > tol) against the snapshot, fields in the order
        isrc * 12 + itf - 1."""
        n = C.c_int()
        _check(self._lib.fmt2d_field_compare(self._h, float(tol), None, None, None, 0, C.byref(n)))
        m = np.zeros(n.value)
        c = np.zeros(n.value, dtype=np.int64)
        o = np.zeros(n.value, dtype=np.int64)
        _llp = C.POINTER(C.c_longlong)
        _check(self._lib.fmt2d_field_compare(self._h, float(tol), _p(m), c.ctypes.data_as(_llp),
                                             o.ctypes.data_as(_llp), n.value, C.byref(n)))
        return m, c, o

    def set_lsqr_mode(self, mode: int = 0):
        """0: LSQR reductions in the reference's order (bit-exact); 1: tree /
        wave reductions."""
        _check(self._lib.fmt2d_set_lsqr_mode(self._h, int(mode)))

    def ordered_sumsq(self, x) -> float:
        """Sum of x[i]**2 added front to back in FP64 on the device (the
        ordered LSQR norm kernel)."""
        v = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(1)
        _check(self._lib.fmt2d_ordered_sumsq(self._h, _p(v), len(v), _p(out)))
        return float(out[0])

    def set_launch_cap(self, max_launches: int = 0):
        _check(self._lib.fmt2d_set_launch_cap(self._h, int(max_launches)))

    def set_solver_launch(self, mode: str = "graph"):
        """How the tile solver's launches reach the device: "graph" (batches
        of launches replayed as HIP graphs) or "level" (one cooperative
        dispatch per stage level, its launches separated by grid-wide
        barriers).  Both reach the same fields bit for bit."""
        modes = {"graph": 0, "level": 1}
        if mode not in modes:
            raise ValueError(f"solver launch mode {mode!r}: 'graph' or 'level'")
        _check(self._lib.fmt2d_set_solver_launch(self._h, modes[mode]))

    def stats(self) -> ForwardStats:
        st = ForwardStats()
        _check(self._lib.fmt2d_forward_stats_get(self._h, C.byref(st)))
        return st

    def summary(self):
        """Per-field summary of the last forward (fmt2d_forward_summary), as a
        numpy structured array in the order isrc * 12 + itf - 1."""
        n = C.c_int()
        _check(self._lib.fmt2d_forward_summary(self._h, None, 0, C.byref(n)))
        buf = (FieldSummary * n.value)()
        _check(self._lib.fmt2d_forward_summary(self._h, buf, n.value, C.byref(n)))
        dt = np.dtype([("reached", np.int64), ("hash", np.uint64), ("max_resid", np.float64),
                       ("mean_resid", np.float64), ("n_resid", np.int64), ("n_resid_big", np.int64)])
        return np.frombuffer(bytes(buf), dtype=dt, count=n.value).copy()

    def forward_trace(self):
        """(due tiles, most sweeps of a tile visit) per launch of the last forward."""
        n = C.c_int()
        _check(self._lib.fmt2d_forward_trace(self._h, None, None, 0, C.byref(n)))
        due = np.zeros(max(n.value, 1), dtype=np.int32)
        mx = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(self._lib.fmt2d_forward_trace(self._h, due.ctypes.data_as(_ip), mx.ctypes.data_as(_ip), n.value,
                                             C.byref(n)))
        return due[:n.value], mx[:n.value]

    def sweep_hist(self):
        """Tile visits of the last forward by sweep count (index k: k + 1 sweeps)."""
        h = np.zeros(64, dtype=np.int64)
        _check(self._lib.fmt2d_forward_sweep_hist(self._h, h.ctypes.data_as(C.POINTER(C.c_longlong))))
        return h

    def set_inversion(self, invert_velocity: bool = True, ifaces: Sequence[int] = ()):
        ids = np.ascontiguousarray(ifaces, dtype=np.int32)
        _check(self._lib.fmt2d_set_inversion(self._h, int(bool(invert_velocity)), len(ids),
                                             ids.ctypes.data_as(_ip) if len(ids) else None))

    def set_inversion_mesh(self, vel_stride: int = 1, iface_stride: int = 1):
        """Inversion mesh: velocity parameters at every vel_stride-th grid
        column and row, interface depths at every iface_stride-th node."""
        _check(self._lib.fmt2d_set_inversion_mesh(self._h, int(vel_stride), int(iface_stride)))

    def set_frechet_chunk(self, max_summands: int = 1 << 28):
        """Contributions sorted at once when the Frechet rows are assembled."""
        _check(self._lib.fmt2d_set_frechet_chunk(self._h, int(max_summands)))

    def inversion_info(self):
        """(velocity parameters, all parameters) of the inversion mesh."""
        a, b = C.c_int(), C.c_int()
        _check(self._lib.fmt2d_inversion_info(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def mesh_info(self):
        """(mesh velocity columns, rows, number of Frechet columns)."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(self._lib.fmt2d_mesh_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def trace(self, src, path, rx, rz):
        """Rays of the last forward: returns (receiver times, valid flags)."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        path = np.ascontiguousarray(path, dtype=np.int32)
        rx = np.ascontiguousarray(rx, dtype=np.float64)
        rz = np.ascontiguousarray(rz, dtype=np.float64)
        n = len(src)
        if not (len(path) == len(rx) == len(rz) == n):
            raise ValueError("src, path, rx and rz must have the same length")
        t = np.zeros(n)
        v = np.zeros(n, dtype=np.int32)
        self._nray = 0
        _check(self._lib.fmt2d_trace(self._h, n, src.ctypes.data_as(_ip), path.ctypes.data_as(_ip), _p(rx), _p(rz),
                                     _p(t), v.ctypes.data_as(_ip)))
        self._nray = n
        self._last_trace_args = (src, path, rx, rz)
        return t, v.astype(bool)

    def set_ray_search(self, mode: int = 0):
        """0: indexed nearest-node search (default); 1: the full scan."""
        _check(self._lib.fmt2d_set_ray_search(self._h, int(mode)))

    def ray_nearest_calls(self) -> int:
        n = C.c_longlong()
        _check(self._lib.fmt2d_ray_nearest_calls(self._h, C.byref(n)))
        return n.value

    def ray_nearest_check(self, region: int, px, pz):
        """(column-index node, scan node) per point of region `region`."""
        x = np.ascontiguousarray(px, dtype=np.float64)
        z = np.ascontiguousarray(pz, dtype=np.float64)
        if len(x) != len(z):
            raise ValueError("px and pz must have the same length")
        a = np.zeros(len(x), dtype=np.int32)
        b = np.zeros(len(x), dtype=np.int32)
        _check(self._lib.fmt2d_ray_nearest_check(self._h, int(region), len(x), _p(x), _p(z), a.ctypes.data_as(_ip),
                                                 b.ctypes.data_as(_ip)))
        return a, b

    def ray_info(self, iray: int):
        e = np.zeros(2)
        w = C.c_double()
        npts = np.zeros(4, dtype=np.int32)
        fl = np.zeros(3, dtype=np.int32)
        _check(self._lib.fmt2d_ray_info(self._h, iray, _p(e), C.byref(w), npts.ctypes.data_as(_ip),
                                        fl.ctypes.data_as(_ip)))
        return {"entry": e, "water_time": w.value, "npts": npts, "headwave": bool(fl[0]), "diffracted": bool(fl[1]),
                "direct": bool(fl[2])}

    def ray_points(self, iray: int, section: int):
        n = C.c_int()
        _check(self._lib.fmt2d_ray_points(self._h, iray, section, None, None, 0, C.byref(n)))
        x, z = np.zeros(max(n.value, 1)), np.zeros(max(n.value, 1))
        _check(self._lib.fmt2d_ray_points(self._h, iray, section, _p(x), _p(z), len(x), C.byref(n)))
        return x[:n.value], z[:n.value]

    def frechet_apply(self, x, transpose: bool = False):
        """G x (x per parameter) or G^T x (x per ray of the last trace) with the
        unscaled Frechet rows, sums in the reference's order."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n_par = self.mesh_info()[2]
        nray = self._nray
        if len(x) != (nray if transpose else n_par):
            raise ValueError("x has the wrong length")
        y = np.zeros(n_par if transpose else nray)
        _check(self._lib.fmt2d_frechet_apply(self._h, int(bool(transpose)), _p(x), _p(y)))
        return y

    def frechet_apply_noalloc(self, x, y, transpose: bool = False):
        """frechet_apply into the caller's y without device allocation (the
        rows must have been used once by frechet_apply on this trace)."""
        n_par = self.mesh_info()[2]
        nray = self._nray
        if not (x.dtype == np.float64 and y.dtype == np.float64 and x.flags.c_contiguous and y.flags.c_contiguous):
            raise TypeError("x and y must be contiguous float64 arrays")
        if len(x) != (nray if transpose else n_par) or len(y) != (n_par if transpose else nray):
            raise ValueError("x or y has the wrong length")
        _check(self._lib.fmt2d_frechet_apply_noalloc(self._h, int(bool(transpose)), _p(x), _p(y)))
        return y

    def adjoint_gradient(self, dJdt):
        """Adjoint-state gradient of a receiver-time misfit (fmt2d_adjoint_gradient):
        dJdt per ray of the last trace; returns (d J / d velocity on the
        propagation grid (nz, nx), d J / d depth of the interface nodes
        (5, nx): interfaces 1..5, AdjointStats)."""
        g = np.ascontiguousarray(np.nan_to_num(np.asarray(dJdt, dtype=np.float64), nan=0.0))
        if len(g) != self._nray:
            raise ValueError(f"dJdt has {len(g)} entries for {self._nray} rays of the last trace")
        info = self.info
        gv = np.zeros(info.nz * info.nx)
        gz = np.zeros(5 * info.nx)
        st = AdjointStats()
        _check(self._lib.fmt2d_adjoint_gradient(self._h, _p(g), _p(gv), _p(gz), C.byref(st)))
        return gv.reshape(info.nz, info.nx), gz.reshape(5, info.nx), st

    def set_adjoint_memory(self, nbytes: int = 0):
        """Cap on the adjoint sweep's edge records per batch (0: 4 GiB)."""
        _check(self._lib.fmt2d_set_adjoint_memory(self._h, int(nbytes)))

    def param_index(self):
        """Frechet column (1-based, 0 none) of each velocity-grid node, (nz, nx);
        the inversion parameters when the inversion mesh is the grid."""
        n = C.c_int()
        _check(self._lib.fmt2d_param_index(self._h, None, 0, C.byref(n)))
        idx = np.zeros(n.value, dtype=np.int32)
        _check(self._lib.fmt2d_param_index(self._h, idx.ctypes.data_as(_ip), len(idx), C.byref(n)))
        return idx.reshape(self.info.nz, self.info.nx)

    def frechet_csr(self, index_base: int = 0):
        """(row_ptr, col, val) of the Frechet rows of the last trace."""
        nnz = C.c_int()
        nr = self._nray_last()
        rp = np.zeros(nr + 1, dtype=np.int32)
        _check(self._lib.fmt2d_frechet_csr(self._h, index_base, rp.ctypes.data_as(_ip), None, None, 0, C.byref(nnz)))
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        val = np.zeros(max(nnz.value, 1))
        _check(self._lib.fmt2d_frechet_csr(self._h, index_base, rp.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                                           _p(val), len(col), C.byref(nnz)))
        return rp, col[:nnz.value], val[:nnz.value]

    def build_system(self, t_obs, crit_chi=1.0, phase_wt_psed=1.0, phase_wt_pmp=1.0, phase_wt_pgss=1.0,
                     refl_weight=10.0, alt_mode=0):
        t = np.ascontiguousarray(t_obs, dtype=np.float64)
        if len(t) != self._nray:
            raise ValueError(f"t_obs has {len(t)} entries for {self._nray} rays of the last trace")
        p = SystemParams(crit_chi, phase_wt_psed, phase_wt_pmp, phase_wt_pgss, refl_weight, alt_mode)
        nd, nz, eps = C.c_int(), C.c_int(), C.c_double()
        _check(self._lib.fmt2d_build_system(self._h, _p(t), C.byref(p), C.byref(nd), C.byref(nz), C.byref(eps)))
        self._ndata = nd.value
        return nd.value, nz.value, eps.value

    def system(self):
        nv, npar = self.inversion_info()
        g = np.zeros(npar)
        r = np.zeros(max(self._ndata, 1))
        _check(self._lib.fmt2d_system_get(self._h, _p(g), _p(r)))
        return g, r[:self._ndata]

    def lsqr(self, dm_sum=None, wdv=0.1, wdv_iface=0.1, w_smooth=0.3, w_smooth_iface=0.3, smooth_type=1,
             max_iter=400, corr_x=0.0, corr_z=0.0, corr_x_iface=0.0):
        """One LSQR solve of the built system; returns (model update, LsqrStats)."""
        nv, npar = self.inversion_info()
        dms = np.zeros(npar) if dm_sum is None else np.ascontiguousarray(dm_sum, dtype=np.float64)
        out = np.zeros(npar)
        p = LsqrParams(wdv, wdv_iface, w_smooth, w_smooth_iface, smooth_type, max_iter, corr_x, corr_z, corr_x_iface)
        st = LsqrStats()
        _check(self._lib.fmt2d_lsqr(self._h, C.byref(p), _p(dms), _p(out), C.byref(st)))
        return out, st

    def damped_update(self, dm_sum=None, lam_damp=0.1, lam_damp_iface=0.1, target_dv=0.03, target_dd=0.5,
                      max_rel_pert=0.3, w_smooth=0.3, w_smooth_iface=0.3, smooth_type=1, max_iter=400, corr_x=0.0,
                      corr_z=0.0, corr_x_iface=0.0):
        """The iteration's model update with fixed or automatic damping;
        returns (capped update, DampingResult)."""
        nv, npar = self.inversion_info()
        dms = np.zeros(npar) if dm_sum is None else np.ascontiguousarray(dm_sum, dtype=np.float64)
        out = np.zeros(npar)
        p = LsqrParams(0.0, 0.0, w_smooth, w_smooth_iface, smooth_type, max_iter, corr_x, corr_z, corr_x_iface)
        d = DampingParams(lam_damp, lam_damp_iface, target_dv, target_dd, max_rel_pert)
        r = DampingResult()
        _check(self._lib.fmt2d_damped_update(self._h, C.byref(p), C.byref(d), _p(dms), _p(out), C.byref(r)))
        return out, r

    def model(self, what=0):
        """Velocity grid (nz, nx) for what=0, else interface `what`'s node depths."""
        n = C.c_int()
        _check(self._lib.fmt2d_model_get(self._h, int(what), None, 0, C.byref(n)))
        out = np.zeros(n.value)
        _check(self._lib.fmt2d_model_get(self._h, int(what), _p(out), n.value, C.byref(n)))
        if what == 0:
            return out.reshape(self.info.nz, self.info.nx)
        return out

    def model_put(self, what, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        _check(self._lib.fmt2d_model_put(self._h, int(what), _p(v), len(v)))

    def _check_dm(self, dm):
        if len(dm) != self.inversion_info()[1]:
            raise ValueError("dm needs one entry per inversion parameter")

    def apply_model_step(self, step, step_iface, dm):
        dm = np.ascontiguousarray(dm, dtype=np.float64)
        self._check_dm(dm)
        _check(self._lib.fmt2d_apply_model_step(self._h, float(step), float(step_iface), _p(dm)))

    def line_search(self, t_obs, dm, crit_chi=1.0, iface_step_ratio=0.5, first_step=1, valid=None):
        """Line search on the RMS; returns (step-scaled update, LineSearchResult).

        valid: the iteration's ray validity (default: that of the last trace);
        when given, it is updated in place with the rays the trial forwards
        lost."""
        t_obs = np.ascontiguousarray(t_obs, dtype=np.float64)
        if len(t_obs) != self._nray:
            raise ValueError(f"t_obs has {len(t_obs)} entries for {self._nray} rays of the last trace")
        if valid is not None and len(valid) != self._nray:
            raise ValueError(f"valid has {len(valid)} entries for {self._nray} rays of the last trace")
        dm = np.array(dm, dtype=np.float64)
        self._check_dm(dm)
        p = LineSearchParams(crit_chi, iface_step_ratio, first_step)
        r = LineSearchResult()
        vp = None
        if valid is not None:
            if valid.dtype != np.int32 or not valid.flags.c_contiguous:
                raise TypeError("valid must be a contiguous int32 array")
            vp = valid.ctypes.data_as(_ip)
        _check(self._lib.fmt2d_line_search(self._h, C.byref(p), _p(t_obs), _p(dm), vp, C.byref(r)))
        return dm, r

    def _nray_last(self):
        return self._nray

    def field(self, isrc: int, itf: int):
        """(T, dT/dx, dT/dz) of field itf (1..12) of source isrc, reference node order."""
        # region sizes change when the geometry is rebuilt (model steps,
        # interface depths put back): ask for the current ones
        info = GridInfo()
        _check(self._lib.fmt2d_grid_info_get(self._h, C.byref(info)))
        self.info = info
        n = info.region_nnode[self._lib.fmt2d_tf_region(itf)]
        T, gx, gz = np.empty(n), np.empty(n), np.empty(n)
        _check(self._lib.fmt2d_get_field(self._h, isrc, itf, _p(T), _p(gx), _p(gz)))
        return T, gx, gz
