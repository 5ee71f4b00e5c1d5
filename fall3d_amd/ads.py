"""Python host binding of libfall3d_ads_hip.so (the gfx950 ADS step).

Mirrors the public interface of the reference's MODULE ADS
(Sources/mod_ADS.F90:48-55) over the C ABI declared in include/fall3d_ads.h:

  ADS.initialize          <- ADS_initialize(limitertype, time_marching)   :98
  ADS.get_vcomps          <- ADS_get_vcomps(my_u, my_v, my_w)             :166
  ADS.update_horizontal_vcomps <- ADS_update_horizontal_vcomps(...)       :187
  ADS.solve_along_x/y/z   <- ADS_solve_along_x/y/z(...)                   :233/:422/:612
  ADS.solve_1d            <- ADS_solve_1D(...)                            :820
  ADS.release             <- ADS_release()                                :208

plus the resident F3D_time_step (Sources/mod_F3D.F90:1397-1491) methods
(set_meteo, set_source, set_state, step, run, time_step, budget).

Arrays are numpy arrays in the reference layouts (Fortran order, halos as
in the reference; see fall3d_amd/problem.py). There is no CPU fallback:
constructing an ADS raises if the HIP library is missing or no GPU is
visible.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FALL3D_ADS_LIB") or os.path.join(_HERE, "lib", "libfall3d_ads_hip.so")

MINMOD, SUPERBEE, OSPRE = 1, 2, 3
EULER, RK = 1, 2
STATE_FP32, STATE_FP32X2, STATE_FP64 = 0, 1, 2
LIMITERS = {"minmod": MINMOD, "superbee": SUPERBEE, "ospre": OSPRE}
TIME_MARCHING = {"euler": EULER, "rk": RK, "runge-kutta": RK}
STATES = {"fp32": STATE_FP32, "fp32x2": STATE_FP32X2, "fp64": STATE_FP64}
BOUNDARIES = {"freeflow": 0, "dirichlet": 1, "periodic": 2}

BUDGET_MASS_VOLUME, BUDGET_MASS_CELLS, BUDGET_GROUND, BUDGET_LATERAL, BUDGET_INJECTED, BUDGET_LEDGER = 0, 1, 2, 3, 4, 5
BUDGET_LEN = 12
KERNELS = {"source": 0, "sweep_x": 1, "sweep_y": 2, "sweep_z": 3, "reduce": 4}

# Every entry point the header declares (checked by the CPU tests).
EXPORTS = (
    "fall3d_ads_initialize", "fall3d_ads_set_grid", "fall3d_ads_get_vcomps",
    "fall3d_ads_update_horizontal_vcomps", "fall3d_ads_solve_along_x", "fall3d_ads_solve_along_y",
    "fall3d_ads_solve_along_z", "fall3d_ads_solve_1d", "fall3d_ads_release", "fall3d_ads_last_error",
    "fall3d_ads_set_meteo", "fall3d_ads_set_source", "fall3d_ads_set_source_metric", "fall3d_ads_set_state", "fall3d_ads_get_state",
    "fall3d_ads_get_time_step", "fall3d_ads_step", "fall3d_ads_step_noalloc", "fall3d_ads_run", "fall3d_ads_get_fluxes",
    "fall3d_ads_set_fluxes", "fall3d_ads_get_budget", "fall3d_ads_set_stream", "fall3d_ads_synchronize",
    "fall3d_ads_set_tiling", "fall3d_ads_set_profiling", "fall3d_ads_get_kernel_time", "fall3d_ads_sweep_bytes",
    "fall3d_ads_copy_probe", "fall3d_ads_copy_probe64", "fall3d_ads_check_state", "fall3d_ads_learned_fallbacks", "fall3d_ads_snapshot",
    "fall3d_ads_restore", "fall3d_ens_statistics", "fall3d_ads_update_meteo", "fall3d_ads_get_meteo",
    "fall3d_ads_postprocess", "fall3d_ads_postprocess_cuts", "fall3d_ads_set_boundary",
    "fall3d_ads_set_graph", "fall3d_ads_autotune", "fall3d_ads_get_tiling",
    "fall3d_ads_ensemble_create", "fall3d_ads_ensemble_release", "fall3d_ads_ensemble_last_error",
    "fall3d_ads_ensemble_member", "fall3d_ads_ensemble_step", "fall3d_ads_ensemble_run",
    "fall3d_ads_ensemble_set_graph", "fall3d_ads_ensemble_synchronize", "fall3d_ads_ensemble_set_tiling",
    "fall3d_ads_ensemble_get_tiling", "fall3d_ads_ensemble_autotune",
    "fall3d_mlp_create", "fall3d_mlp_release", "fall3d_mlp_last_error", "fall3d_mlp_nparam",
    "fall3d_mlp_set_params", "fall3d_mlp_get_params", "fall3d_mlp_forward", "fall3d_mlp_gradient",
    "fall3d_mlp_data_append", "fall3d_mlp_data_count", "fall3d_mlp_data_get", "fall3d_mlp_data_clear",
    "fall3d_mlp_data_refs",
    "fall3d_mlp_train", "fall3d_ads_set_flux_output", "fall3d_ads_get_face_flux", "fall3d_ads_set_learned",
    "fall3d_ads_coarse_grain", "fall3d_ads_learned_collect",
)


class Fall3dAdsError(RuntimeError):
    pass


# fall3d_meteo_options / fall3d_meteo_grid / fall3d_meteo_corner of include/fall3d_ads.h
class _MeteoOptions(ctypes.Structure):
    _fields_ = [("map_h", ctypes.c_int), ("map_v", ctypes.c_int), ("modkh", ctypes.c_int), ("modkv", ctypes.c_int),
                ("modv", ctypes.c_int), ("nbins", ctypes.c_int), ("kh0", ctypes.c_double), ("kv0", ctypes.c_double),
                ("rams_cs", ctypes.c_double), ("dlon", ctypes.c_double), ("dlat", ctypes.c_double),
                ("x3max", ctypes.c_double), ("s_time", ctypes.c_double), ("vset_fac", ctypes.c_double),
                ("diam", ctypes.c_void_p), ("rho", ctypes.c_void_p), ("psi", ctypes.c_void_p),
                ("cat", ctypes.c_void_p), ("aggregate", ctypes.c_void_p), ("dry_deposition", ctypes.c_int)]


class _MeteoGrid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("Hm1_p", "Hm2_p", "Hm3_p", "dhdx_p", "dhdy_p", "z_c", "h_c", "X3_c",
                                               "X3_p")]


class _MeteoCorner(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p * 2) for n in ("uc", "vc", "wc", "tc", "tvc", "rhoc", "pblhc", "ustc", "monc")] + \
        [("lusec", ctypes.c_void_p)]


class _OutputArgs(ctypes.Structure):
    _fields_ = [("Hm1_p", ctypes.c_void_p), ("Hm2_p", ctypes.c_void_p), ("Hm3_p", ctypes.c_void_p),
                ("z_c", ctypes.c_void_p), ("bin_species", ctypes.c_void_p), ("species", ctypes.c_int),
                ("species_category", ctypes.c_int), ("so2_du", ctypes.c_int), ("npm", ctypes.c_int),
                ("pm_um", ctypes.c_void_p), ("bin_diam", ctypes.c_void_p)]


class _CutArgs(ctypes.Structure):
    _fields_ = [("lonmin", ctypes.c_double), ("dlon", ctypes.c_double), ("latmin", ctypes.c_double),
                ("dlat", ctypes.c_double), ("sigma", ctypes.c_void_p), ("X3max", ctypes.c_double),
                ("ncutx", ctypes.c_int), ("ncuty", ctypes.c_int), ("ncutz", ctypes.c_int), ("nfl", ctypes.c_int),
                ("x_cut", ctypes.c_void_p), ("y_cut", ctypes.c_void_p), ("z_cut", ctypes.c_void_p),
                ("fl_cut", ctypes.c_void_p)]


METEO_MAP_H = {"cartesian": 0, "spherical": 1}
METEO_MAP_V = {"cartesian": 0, "sigma": 1, "sigma_linear_decay": 2}
METEO_KH = {"constant": 1, "cmaq": 2, "rams": 4}
METEO_KV = {"constant": 1, "similarity": 3}
METEO_VSET = {"arastopour": 1, "ganser": 2, "wilson": 3, "dellino": 4, "pfeiffer": 5, "dioguardi2017": 6,
              "dioguardi2018": 7}


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_int = ctypes.c_int
_dbl = ctypes.c_double


def load_library(path: str = LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not found: build it with `python -m fall3d_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and
    # the dynamic loader binds this library to whichever copy of that soname
    # is loaded first. Loading torch first keeps torch's device use working
    # in a process that also calls this library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    sig = {
        "fall3d_ads_initialize": ([ctypes.POINTER(_vp), _int, _int, _int, _int, _int, _int, _int, _int], _int),
        "fall3d_ads_set_grid": ([_vp] * 7, _int),
        "fall3d_ads_get_vcomps": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)], _int),
        "fall3d_ads_update_horizontal_vcomps": ([_vp, _vp, _vp, _vp, _vp, _dbl], _int),
        "fall3d_ads_solve_along_x": ([_vp, _dp, _dp, _dbl, _vp, _vp], _int),
        "fall3d_ads_solve_along_y": ([_vp, _dp, _dp, _dbl, _vp, _vp], _int),
        "fall3d_ads_solve_along_z": ([_vp, _dp, _dp, _dbl, _vp, _vp, _vp, _vp, _vp, _dbl, _vp], _int),
        "fall3d_ads_solve_1d": ([_vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dp, _dp, ctypes.POINTER(_int), _int, _int,
                                 _int], _int),
        "fall3d_ads_release": ([_vp], _int),
        "fall3d_ads_last_error": ([_vp], ctypes.c_char_p),
        "fall3d_ads_set_meteo": ([_vp] * 11, _int),
        "fall3d_ads_set_source": ([_vp, _vp], _int),
        "fall3d_ads_set_source_metric": ([_vp, _vp, _vp, _vp, _vp], _int),
        "fall3d_ads_set_state": ([_vp, _vp, _vp], _int),
        "fall3d_ads_get_state": ([_vp, _vp, _vp], _int),
        "fall3d_ads_get_time_step": ([_vp, _dbl, _int, _dbl, _dp], _int),
        "fall3d_ads_step": ([_vp, _dbl, _dbl, _int], _int),
        "fall3d_ads_step_noalloc": ([_vp, _dbl, _dbl, _int], _int),
        "fall3d_ads_copy_probe": ([_vp, _vp, ctypes.c_size_t, _vp], _int),
        "fall3d_ads_copy_probe64": ([_vp, _vp, ctypes.c_size_t, _vp], _int),
        "fall3d_ads_check_state": ([_vp, _vp, _dp], _int),
        "fall3d_ads_snapshot": ([_vp, _vp, ctypes.POINTER(ctypes.c_size_t), _dbl, ctypes.c_longlong], _int),
        "fall3d_ads_restore": ([_vp, _vp, ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_longlong)], _int),
        "fall3d_ens_statistics": ([_vp, _int, ctypes.c_long, _vp, _int, _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _vp], _int),
        "fall3d_ads_learned_fallbacks": ([_vp, _vp], _int),
        "fall3d_ads_update_meteo": ([_vp, _vp, _vp, _vp], _int),
        "fall3d_ads_get_meteo": ([_vp] * 11, _int),
        "fall3d_ads_postprocess": ([_vp] * 10, _int),
        "fall3d_ads_postprocess_cuts": ([_vp] * 7, _int),
        "fall3d_ads_set_boundary": ([_vp, _int], _int),
        "fall3d_ads_run": ([_vp, _dbl, _dbl, _int, _int], _int),
        "fall3d_ads_get_fluxes": ([_vp, _dp], _int),
        "fall3d_ads_set_fluxes": ([_vp, _dp], _int),
        "fall3d_ads_get_budget": ([_vp, _dp], _int),
        "fall3d_ads_set_stream": ([_vp, _vp], _int),
        "fall3d_ads_synchronize": ([_vp], _int),
        "fall3d_ads_set_tiling": ([_vp, _int, _int, _int, _int], _int),
        "fall3d_ads_set_profiling": ([_vp, _int], _int),
        "fall3d_ads_get_kernel_time": ([_vp, _int, _dp, ctypes.POINTER(ctypes.c_longlong)], _int),
        "fall3d_ads_sweep_bytes": ([_vp, _int, _dp], _int),
        "fall3d_ads_set_graph": ([_vp, _int], _int),
        "fall3d_ads_autotune": ([_vp, _dbl, _dbl, _int, _vp, _vp], _int),
        "fall3d_ads_get_tiling": ([_vp, _int, ctypes.POINTER(_int), ctypes.POINTER(_int)], _int),
        "fall3d_ads_ensemble_create": ([ctypes.POINTER(_vp), _int, ctypes.POINTER(_int), _int, _int, _int, _int,
                                        _int], _int),
        "fall3d_ads_ensemble_release": ([_vp], _int),
        "fall3d_ads_ensemble_last_error": ([_vp], ctypes.c_char_p),
        "fall3d_ads_ensemble_member": ([_vp, _int, ctypes.POINTER(_vp)], _int),
        "fall3d_ads_ensemble_step": ([_vp, _dp, _dp, _int], _int),
        "fall3d_ads_ensemble_run": ([_vp, _dp, _dp, _int, _int], _int),
        "fall3d_ads_ensemble_set_graph": ([_vp, _int], _int),
        "fall3d_ads_ensemble_synchronize": ([_vp], _int),
        "fall3d_ads_ensemble_set_tiling": ([_vp, _int, _int, _int], _int),
        "fall3d_ads_ensemble_get_tiling": ([_vp, ctypes.POINTER(_int)], _int),
        "fall3d_ads_ensemble_autotune": ([_vp, _dp, _dp, _int, _dp], _int),
        "fall3d_mlp_create": ([ctypes.POINTER(_vp), ctypes.c_long], _int),
        "fall3d_mlp_release": ([_vp], _int),
        "fall3d_mlp_last_error": ([_vp], ctypes.c_char_p),
        "fall3d_mlp_nparam": ([], _int),
        "fall3d_mlp_set_params": ([_vp, _vp], _int),
        "fall3d_mlp_get_params": ([_vp, _vp], _int),
        "fall3d_mlp_forward": ([_vp, _vp, ctypes.c_long, _vp], _int),
        "fall3d_mlp_gradient": ([_vp, _vp, _vp, ctypes.c_long, _vp, _dp], _int),
        "fall3d_mlp_data_append": ([_vp, _vp, _vp, ctypes.c_long], _int),
        "fall3d_mlp_data_count": ([_vp, ctypes.POINTER(ctypes.c_long)], _int),
        "fall3d_mlp_data_get": ([_vp, ctypes.c_long, ctypes.c_long, _vp, _vp], _int),
        "fall3d_mlp_data_clear": ([_vp], _int),
        "fall3d_mlp_data_refs": ([_vp, ctypes.c_long, ctypes.c_long, _vp, _vp], _int),
        "fall3d_mlp_train": ([_vp, _int, _int, ctypes.c_float, ctypes.c_ulonglong, _vp], _int),
        "fall3d_ads_set_flux_output": ([_vp, _int], _int),
        "fall3d_ads_get_face_flux": ([_vp, _int, _vp], _int),
        "fall3d_ads_set_learned": ([_vp, _vp, _dbl, _dbl, _dbl], _int),
        "fall3d_ads_coarse_grain": ([_vp, _vp], _int),
        "fall3d_ads_learned_collect": ([_vp, _vp, _vp, _dbl, _dbl, _int, _dbl, ctypes.POINTER(ctypes.c_long)],
                                       _int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _kind_dtype(kind):
    return np.float64 if kind == 8 else np.float32


class ADS:
    """One ADS instance (a context of libfall3d_ads_hip.so) for an nx*ny*nz*nbins domain."""

    def __init__(self, nx, ny, nz, nbins, limiter="minmod", time_marching="rk", real_kind=8, state="fp32",
                 grid=None, _borrow=None):
        L = load_library()
        self._borrowed = _borrow is not None
        self.nx, self.ny, self.nz, self.nb = int(nx), int(ny), int(nz), int(nbins)
        self.kind = int(real_kind)
        self.dtype = _kind_dtype(self.kind)
        lim = LIMITERS[limiter] if isinstance(limiter, str) else int(limiter)
        tm = TIME_MARCHING[time_marching] if isinstance(time_marching, str) else int(time_marching)
        st = STATES[state] if isinstance(state, str) else int(state)
        self.limiter, self.time_marching, self.state = lim, tm, st
        self._keep = {}
        if _borrow is not None:  # a member context owned by an Ensemble
            self._h = _borrow
            return
        h = _vp()
        rc = L.fall3d_ads_initialize(ctypes.byref(h), self.nx, self.ny, self.nz, self.nb, lim, tm, self.kind, st)
        self._h = h
        if rc != 0:
            msg = L.fall3d_ads_last_error(h).decode()
            L.fall3d_ads_release(h)
            self._h = None
            raise Fall3dAdsError(f"fall3d_ads_initialize failed ({rc}): {msg}")
        if grid is not None:
            self.set_grid(grid)

    # -- helpers -------------------------------------------------------------
    def _chk(self, rc, what):
        if rc != 0:
            msg = load_library().fall3d_ads_last_error(self._h).decode()
            raise Fall3dAdsError(f"{what} failed ({rc}): {msg}")

    def _arr(self, a, shape, name, writable=False):
        if a is None:
            return None
        a = np.asarray(a)
        if a.dtype != self.dtype:
            if writable:
                raise TypeError(f"{name}: expected dtype {self.dtype}, got {a.dtype}")
            a = a.astype(self.dtype)
        if a.size != int(np.prod(shape)):
            raise ValueError(f"{name}: expected {shape} elements, got shape {a.shape}")
        # a strided view (also a 1-D one such as x[::2]) would hand the
        # library the wrong elements: read-only inputs are copied, writable
        # arguments must already be contiguous
        if not a.flags.f_contiguous:
            if writable:
                raise ValueError(f"{name}: must be Fortran-contiguous")
            a = np.asfortranarray(a)
        return a

    def _p(self, a):
        return None if a is None else a.ctypes.data

    # -- shapes (reference layouts) -------------------------------------------
    @property
    def shape_c(self):
        return (self.nx + 4, self.ny + 4, self.nz + 4, self.nb)

    @property
    def shape_acum(self):
        return (self.nx + 4, self.ny + 4, self.nb)

    @property
    def shape_u(self):
        return (self.nx + 3, self.ny, self.nz)

    @property
    def shape_v(self):
        return (self.ny + 3, self.nx, self.nz)

    @property
    def shape_w(self):
        return (self.nz + 3, self.nx, self.ny)

    @property
    def shape_vs(self):
        return (self.nx, self.ny, self.nz + 3, self.nb)

    # -- MODULE ADS --------------------------------------------------------------
    def set_grid(self, grid):
        """ARAKAWA_C_GRID spacings dX1_p, dX1_b, dX2_p, dX2_b, dX3_p, dX3_b (dict)."""
        arrs = [np.ascontiguousarray(grid[k], dtype=self.dtype)
                for k in ("dX1_p", "dX1_b", "dX2_p", "dX2_b", "dX3_p", "dX3_b")]
        exp = [self.nx + 4, self.nx + 3, self.ny + 4, self.ny + 3, self.nz + 4, self.nz + 3]
        for a, n in zip(arrs, exp):
            if a.size != n:
                raise ValueError(f"grid array of size {a.size}, expected {n}")
        self._chk(load_library().fall3d_ads_set_grid(self._h, *[a.ctypes.data for a in arrs]), "set_grid")
        self._keep["grid"] = arrs

    def get_vcomps(self):
        """Host arrays CB%u, CB%v, CB%w (writable numpy views owned by the context)."""
        u, v, w = _vp(), _vp(), _vp()
        self._chk(load_library().fall3d_ads_get_vcomps(self._h, ctypes.byref(u), ctypes.byref(v), ctypes.byref(w)),
                  "get_vcomps")
        ct = ctypes.c_double if self.kind == 8 else ctypes.c_float

        def view(ptr, shape):
            n = int(np.prod(shape))
            buf = (ct * n).from_address(ptr.value)
            return np.frombuffer(buf, dtype=self.dtype).reshape(shape, order="F")

        return view(u, self.shape_u), view(v, self.shape_v), view(w, self.shape_w)

    def update_horizontal_vcomps(self, u1, u2, v1, v2, stime):
        u1 = self._arr(u1, self.shape_u, "u1"); u2 = self._arr(u2, self.shape_u, "u2")
        v1 = self._arr(v1, self.shape_v, "v1"); v2 = self._arr(v2, self.shape_v, "v2")
        self._chk(load_library().fall3d_ads_update_horizontal_vcomps(self._h, u1.ctypes.data, u2.ctypes.data,
                                                                     v1.ctypes.data, v2.ctypes.data, float(stime)),
                  "update_horizontal_vcomps")

    def solve_along_x(self, W_flux, E_flux, dt, c, k1):
        c = self._arr(c, self.shape_c, "c", writable=True)
        k1 = self._arr(k1, (self.nx + 4, self.ny, self.nz), "k1")
        W, E = _dbl(W_flux), _dbl(E_flux)
        self._chk(load_library().fall3d_ads_solve_along_x(self._h, ctypes.byref(W), ctypes.byref(E), float(dt),
                                                          c.ctypes.data, k1.ctypes.data), "solve_along_x")
        return W.value, E.value

    def solve_along_y(self, S_flux, N_flux, dt, c, k2):
        c = self._arr(c, self.shape_c, "c", writable=True)
        k2 = self._arr(k2, (self.ny + 4, self.nx, self.nz), "k2")
        S, N = _dbl(S_flux), _dbl(N_flux)
        self._chk(load_library().fall3d_ads_solve_along_y(self._h, ctypes.byref(S), ctypes.byref(N), float(dt),
                                                          c.ctypes.data, k2.ctypes.data), "solve_along_y")
        return S.value, N.value

    def solve_along_z(self, D_flux, U_flux, dt, c, acum, k3, w1, w2, stime, vs):
        c = self._arr(c, self.shape_c, "c", writable=True)
        acum = self._arr(acum, self.shape_acum, "acum", writable=True)
        k3 = self._arr(k3, (self.nz + 4, self.nx, self.ny), "k3")
        w1 = self._arr(w1, self.shape_w, "w1")
        w2 = self._arr(w2, self.shape_w, "w2")
        vs = self._arr(vs, self.shape_vs, "vs")
        D, U = _dbl(D_flux), _dbl(U_flux)
        self._chk(load_library().fall3d_ads_solve_along_z(self._h, ctypes.byref(D), ctypes.byref(U), float(dt),
                                                          c.ctypes.data, acum.ctypes.data, k3.ctypes.data,
                                                          w1.ctypes.data, w2.ctypes.data, float(stime),
                                                          vs.ctypes.data), "solve_along_z")
        return D.value, U.value

    def solve_1d(self, c, dxp, dxb, u, k, dt, bvalue, bcode, ips, ipe, limiter):
        """ADS_solve_1D on one line; c is updated in place, returns flux(2)."""
        n = ipe - ips + 1
        c = self._arr(c, (n + 4,), "c", writable=True)
        dxp = self._arr(dxp, (n + 4,), "dxp"); k = self._arr(k, (n + 4,), "k")
        dxb = self._arr(dxb, (n + 3,), "dxb"); u = self._arr(u, (n + 3,), "u")
        bv = (ctypes.c_double * 2)(*bvalue)
        fl = (ctypes.c_double * 2)()
        bc = (ctypes.c_int * 2)(*bcode)
        self._chk(load_library().fall3d_ads_solve_1d(self._h, c.ctypes.data, dxp.ctypes.data, dxb.ctypes.data,
                                                     u.ctypes.data, k.ctypes.data, float(dt), bv, fl, bc, int(ips),
                                                     int(ipe), int(limiter)), "solve_1d")
        return fl[0], fl[1]

    def release(self):
        if getattr(self, "_h", None):
            if not self._borrowed:
                load_library().fall3d_ads_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    # -- resident F3D_time_step --------------------------------------------------
    def set_meteo(self, u1, v1, w1, k1, k2, k3, vs, u2=None, v2=None, w2=None):
        a = {
            "u1": self._arr(u1, self.shape_u, "u1"), "u2": self._arr(u2, self.shape_u, "u2"),
            "v1": self._arr(v1, self.shape_v, "v1"), "v2": self._arr(v2, self.shape_v, "v2"),
            "w1": self._arr(w1, self.shape_w, "w1"), "w2": self._arr(w2, self.shape_w, "w2"),
            "k1": self._arr(k1, (self.nx + 4, self.ny, self.nz), "k1"),
            "k2": self._arr(k2, (self.ny + 4, self.nx, self.nz), "k2"),
            "k3": self._arr(k3, (self.nz + 4, self.nx, self.ny), "k3"),
            "vs": self._arr(vs, self.shape_vs, "vs"),
        }
        p = self._p
        self._chk(load_library().fall3d_ads_set_meteo(self._h, p(a["u1"]), p(a["u2"]), p(a["v1"]), p(a["v2"]),
                                                      p(a["w1"]), p(a["w2"]), p(a["k1"]), p(a["k2"]), p(a["k3"]),
                                                      p(a["vs"])), "set_meteo")

    def update_meteo(self, corner, grid, opts):
        """F3D_update_meteo_term on the device (fall3d_ads_update_meteo).

        corner: dict of the corner fields, each a pair (first time, second
        time): "uc", "vc", "wc", "tc", "rhoc" of shape (nx+1, ny+1, nz+1) and,
        for similarity, "tvc" and the 2-D "pblhc", "ustc", "monc" (nx+1, ny+1);
        for dry deposition "ustc" and the land use "lusec" (one 2-D array).
        grid: "Hm1_p" (ny), "Hm2_p" (ny+4), "Hm3_p" (nx, ny), and as the
        options need them "dhdx_p", "dhdy_p", "z_c", "h_c", "X3_c", "X3_p".
        opts: "map_h", "map_v", "modkh", "modkv", "modv" (codes or names of
        METEO_*), "kh0", "kv0", "rams_cs", "dlon", "dlat", "x3max", "s_time",
        "vset_fac", the per-bin "diam", "rho", "psi", "cat", "aggregate" and
        "dry_deposition" (0 or 1)."""
        keep = []

        def arr(x, shape, name):
            if x is None:
                return None
            a = self._arr(x, shape, name)
            keep.append(a)
            return a.ctypes.data

        nx, ny, nz, nb = self.nx, self.ny, self.nz, self.nb
        c3, c2 = (nx + 1, ny + 1, nz + 1), (nx + 1, ny + 1)
        cm = _MeteoCorner()
        for key, shp in (("uc", c3), ("vc", c3), ("wc", c3), ("tc", c3), ("tvc", c3), ("rhoc", c3),
                         ("pblhc", c2), ("ustc", c2), ("monc", c2)):
            pair = corner.get(key)
            if pair is not None:
                getattr(cm, key)[0] = arr(pair[0], shp, key)
                getattr(cm, key)[1] = arr(pair[1], shp, key)
        cm.lusec = arr(corner.get("lusec"), c2, "lusec")
        gm = _MeteoGrid()
        for key, shp in (("Hm1_p", (ny,)), ("Hm2_p", (ny + 4,)), ("Hm3_p", (nx, ny)), ("dhdx_p", (nx, ny)),
                         ("dhdy_p", (nx, ny)), ("z_c", c3), ("h_c", c2), ("X3_c", (nz + 1,)), ("X3_p", (nz,))):
            setattr(gm, key, arr(grid.get(key), shp, key))
        code = lambda tab, v: tab[v] if isinstance(v, str) else int(v)
        bins = {k: np.ascontiguousarray(opts[k], dtype=np.float64) for k in ("diam", "rho", "psi")}
        bins.update({k: np.ascontiguousarray(opts[k], dtype=np.int32) for k in ("cat", "aggregate")})
        for k, v in bins.items():
            if v.size != nb:
                raise ValueError(f"{k}: {nb} bins expected")
        om = _MeteoOptions(code(METEO_MAP_H, opts.get("map_h", 0)), code(METEO_MAP_V, opts.get("map_v", 0)),
                           code(METEO_KH, opts.get("modkh", 1)), code(METEO_KV, opts.get("modkv", 1)),
                           code(METEO_VSET, opts.get("modv", 2)), nb, float(opts.get("kh0", 0.0)),
                           float(opts.get("kv0", 0.0)), float(opts.get("rams_cs", 0.2)), float(opts.get("dlon", 1.0)),
                           float(opts.get("dlat", 1.0)), float(opts.get("x3max", 0.0)), float(opts.get("s_time", 0.5)),
                           float(opts.get("vset_fac", 1.0)), bins["diam"].ctypes.data, bins["rho"].ctypes.data,
                           bins["psi"].ctypes.data, bins["cat"].ctypes.data, bins["aggregate"].ctypes.data,
                           int(opts.get("dry_deposition", 0)))
        self._chk(load_library().fall3d_ads_update_meteo(self._h, ctypes.byref(cm), ctypes.byref(gm),
                                                         ctypes.byref(om)), "update_meteo")
        del keep

    def get_meteo(self):
        """The context's meteo in the layouts set_meteo takes (host kind)."""
        nx, ny, nz, nb = self.nx, self.ny, self.nz, self.nb
        shapes = {"u1": self.shape_u, "u2": self.shape_u, "v1": self.shape_v, "v2": self.shape_v,
                  "w1": self.shape_w, "w2": self.shape_w, "k1": (nx + 4, ny, nz), "k2": (ny + 4, nx, nz),
                  "k3": (nz + 4, nx, ny), "vs": self.shape_vs}
        out = {k: np.zeros(v, dtype=self.dtype, order="F") for k, v in shapes.items()}
        self._chk(load_library().fall3d_ads_get_meteo(self._h, *(out[k].ctypes.data for k in shapes)), "get_meteo")
        return out

    def postprocess(self, z_c=None, hm1=None, hm2=None, hm3=None, bin_species=None, species=0, species_category=1,
                    so2_du=False, pm_um=(), bin_diam=None, want=("con_bins", "con", "con_sur", "col_mass",
                                                                  "col_mass_pm", "cloud_top", "grn_load",
                                                                  "grn_load_bin")):
        """nc_IO_out_res's result variables from the device state (see
        fall3d_ads_postprocess); returns a dict of FP64 corner arrays."""
        nx, ny, nz, nb = self.nx, self.ny, self.nz, self.nb
        a, keep = self._output_args(z_c, hm1, hm2, hm3, bin_species, species, species_category, so2_du, pm_um,
                                    bin_diam)
        pm = keep[-1]
        c2, c3 = (nx + 1, ny + 1), (nx + 1, ny + 1, nz + 1)
        shapes = {"con_bins": c3 + (nb,), "con": c3, "con_sur": c2, "col_mass": c2,
                  "col_mass_pm": c2 + (max(pm.size, 1),), "cloud_top": c2, "grn_load": c2, "grn_load_bin": c2 + (nb,)}
        out = {k: np.zeros(shapes[k], order="F") for k in want}
        ptr = lambda k: out[k].ctypes.data if k in out else None
        self._chk(load_library().fall3d_ads_postprocess(self._h, ctypes.byref(a), *(ptr(k) for k in shapes)),
                  "postprocess")
        del keep
        return out

    def _output_args(self, z_c, hm1, hm2, hm3, bin_species, species, species_category, so2_du, pm_um, bin_diam):
        """fall3d_output_args and the arrays it points to (kept alive by the
        caller; the last one is the PM threshold array)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        keep = []

        def arr(x, shape, name):
            if x is None:
                return None
            a = self._arr(x, shape, name)
            keep.append(a)
            return a.ctypes.data

        spe = None if bin_species is None else np.ascontiguousarray(bin_species, dtype=np.int32)
        pm = np.ascontiguousarray(pm_um, dtype=np.float64)
        diam = None if bin_diam is None else np.ascontiguousarray(bin_diam, dtype=np.float64)
        a = _OutputArgs(arr(hm1, (ny,), "Hm1_p"), arr(hm2, (ny + 4,), "Hm2_p"), arr(hm3, (nx, ny), "Hm3_p"),
                        arr(z_c, (nx + 1, ny + 1, nz + 1), "z_c"), None if spe is None else spe.ctypes.data,
                        int(species), int(species_category), int(bool(so2_du)), int(pm.size),
                        pm.ctypes.data if pm.size else None, None if diam is None else diam.ctypes.data)
        keep += [spe, diam, pm]
        return a, keep

    def postprocess_cuts(self, z_c, sigma, X3max, lonmin=0.0, dlon=1.0, latmin=0.0, dlat=1.0, x_cut=(), y_cut=(),
                         z_cut=(), fl_cut=(), hm1=None, hm2=None, hm3=None, bin_species=None, species=0):
        """nc_IO_out_res's concentration cuts of one species (see
        fall3d_ads_postprocess_cuts): dict with cut_x (ny+1, nz+1, ncutx),
        cut_y (nx+1, nz+1, ncuty), cut_z (nx+1, ny+1, ncutz) and cut_fl
        (nx+1, ny+1, nfl) in g/m3; fl_cut in flight levels."""
        nx, ny, nz = self.nx, self.ny, self.nz
        a, keep = self._output_args(z_c, hm1, hm2, hm3, bin_species, species, 1, False, (), None)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        if sg.shape != (nz + 1,):
            raise ValueError(f"sigma: expected {nz + 1} corner levels, got shape {sg.shape}")
        pos = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x_cut, y_cut, z_cut, fl_cut)]
        ca = _CutArgs(float(lonmin), float(dlon), float(latmin), float(dlat), sg.ctypes.data, float(X3max),
                      *(int(v.size) for v in pos), *(v.ctypes.data if v.size else None for v in pos))
        shapes = [(ny + 1, nz + 1), (nx + 1, nz + 1), (nx + 1, ny + 1), (nx + 1, ny + 1)]
        out = {k: np.zeros(sh + (v.size,), order="F")
               for k, sh, v in zip(("cut_x", "cut_y", "cut_z", "cut_fl"), shapes, pos)}
        ptrs = [o.ctypes.data if o.size else None for o in out.values()]
        self._chk(load_library().fall3d_ads_postprocess_cuts(self._h, ctypes.byref(a), ctypes.byref(ca), *ptrs),
                  "postprocess_cuts")
        del keep, sg, pos
        return out

    def set_boundary(self, mode="freeflow"):
        """Ghost rule of the sweeps: "freeflow" (default), "dirichlet" or
        "periodic" (the reference's zero-gradient copy)."""
        m = BOUNDARIES[mode] if isinstance(mode, str) else int(mode)
        self._chk(load_library().fall3d_ads_set_boundary(self._h, m), "set_boundary")

    def set_problem(self, prob):
        """Grid, meteo, source and initial state of a fall3d_amd.problem.Problem."""
        self.set_grid(prob.grid)
        self.set_meteo(prob.u1, prob.v1, prob.w1, prob.k1, prob.k2, prob.k3, prob.vs, prob.u2, prob.v2, prob.w2)
        self.set_source(prob.s)
        self.set_state(prob.c0)

    def set_source(self, s, hm1=None, hm2=None, hm3=None):
        """Source rate s(1:nx,1:ny,1:nz,nb); optional map factors Hm1_p(1:ny),
        Hm2_p(-1:ny+2), Hm3_p(1:nx,1:ny) (None = 1, the flat Cartesian grid)."""
        s = self._arr(s, (self.nx, self.ny, self.nz, self.nb), "s") if s is not None else None
        if hm1 is None and hm2 is None and hm3 is None:
            self._chk(load_library().fall3d_ads_set_source(self._h, self._p(s)), "set_source")
            return
        h1 = self._arr(hm1, (self.ny,), "Hm1_p")
        h2 = self._arr(hm2, (self.ny + 4,), "Hm2_p")
        h3 = self._arr(hm3, (self.nx, self.ny), "Hm3_p")
        self._chk(load_library().fall3d_ads_set_source_metric(self._h, self._p(s), self._p(h1), self._p(h2),
                                                              self._p(h3)), "set_source_metric")

    def set_state(self, c, acum=None):
        """Upload c (None: zero field) and acum (None: zero)."""
        c = self._arr(c, self.shape_c, "c") if c is not None else None
        acum = self._arr(acum, self.shape_acum, "acum") if acum is not None else None
        self._chk(load_library().fall3d_ads_set_state(self._h, self._p(c), self._p(acum)), "set_state")

    def get_state(self):
        c = np.zeros(self.shape_c, dtype=self.dtype, order="F")
        acum = np.zeros(self.shape_acum, dtype=self.dtype, order="F")
        self._chk(load_library().fall3d_ads_get_state(self._h, c.ctypes.data, acum.ctypes.data), "get_state")
        return c, acum

    def time_step(self, stime=0.0, criterion=2, cfl=0.9):
        dt = _dbl()
        self._chk(load_library().fall3d_ads_get_time_step(self._h, float(stime), int(criterion), float(cfl),
                                                          ctypes.byref(dt)), "get_time_step")
        return dt.value

    def step(self, dt, iiter, stime=0.0):
        self._chk(load_library().fall3d_ads_step(self._h, float(dt), float(stime), int(iiter)), "step")

    def step_noalloc(self, dt, iiter, stime=0.0):
        """fall3d_ads_step_noalloc: the same step with its partial sums in the
        library's static per-GPU array (no device allocation)."""
        self._chk(load_library().fall3d_ads_step_noalloc(self._h, float(dt), float(stime), int(iiter)),
                  "step_noalloc")

    def run(self, dt, nsteps, iiter0=0, stime=0.0):
        self._chk(load_library().fall3d_ads_run(self._h, float(dt), float(stime), int(iiter0), int(nsteps)), "run")

    def synchronize(self):
        self._chk(load_library().fall3d_ads_synchronize(self._h), "synchronize")

    def fluxes(self):
        f = (ctypes.c_double * 6)()
        self._chk(load_library().fall3d_ads_get_fluxes(self._h, f), "get_fluxes")
        return np.array(f[:])

    def set_fluxes(self, flux):
        f = (ctypes.c_double * 6)(*flux)
        self._chk(load_library().fall3d_ads_set_fluxes(self._h, f), "set_fluxes")

    def budget(self):
        b = (ctypes.c_double * BUDGET_LEN)()
        self._chk(load_library().fall3d_ads_get_budget(self._h, b), "get_budget")
        out = np.array(b[:])
        return {
            "mass_volume": out[0], "mass_cells": out[1], "ground": out[2], "lateral": out[3], "injected": out[4],
            "ledger": out[5:11].copy(), "residual": out[11],
        }

    def snapshot(self, time=0.0, iiter=0):
        """Restart snapshot of the device state and accumulators (bytes)."""
        lib = load_library()
        n = ctypes.c_size_t(0)
        self._chk(lib.fall3d_ads_snapshot(self._h, None, ctypes.byref(n), float(time), int(iiter)), "snapshot")
        buf = np.empty(n.value, dtype=np.uint8)
        self._chk(lib.fall3d_ads_snapshot(self._h, buf.ctypes.data, ctypes.byref(n), float(time), int(iiter)),
                  "snapshot")
        return buf.tobytes()

    def restore(self, blob):
        """Restore a snapshot; returns (time, iiter) stored in it."""
        buf = np.frombuffer(blob, dtype=np.uint8)
        t = ctypes.c_double(0.0)
        it = ctypes.c_longlong(0)
        self._chk(load_library().fall3d_ads_restore(self._h, buf.ctypes.data, buf.size, ctypes.byref(t),
                                                    ctypes.byref(it)), "restore")
        return t.value, int(it.value)

    def save_restart(self, path, time=0.0, iiter=0):
        with open(path, "wb") as f:
            f.write(self.snapshot(time, iiter))

    def load_restart(self, path):
        with open(path, "rb") as f:
            return self.restore(f.read())

    def check_state(self):
        """Non-finite and negative interior cells and the smallest finite value."""
        counts = (ctypes.c_longlong * 2)()
        mn = ctypes.c_double(0.0)
        self._chk(load_library().fall3d_ads_check_state(self._h, counts, ctypes.byref(mn)), "check_state")
        return {"nonfinite": int(counts[0]), "negative": int(counts[1]), "min": mn.value}

    def learned_fallbacks(self):
        """Learned steps that fell back to plain KT (non-finite network output)."""
        n = ctypes.c_longlong(0)
        self._chk(load_library().fall3d_ads_learned_fallbacks(self._h, ctypes.byref(n)), "learned_fallbacks")
        return int(n.value)

    def set_stream(self, stream_handle):
        self._chk(load_library().fall3d_ads_set_stream(self._h, stream_handle), "set_stream")

    def set_tiling(self, axis, cells_per_thread=0, columns=0, bins_per_group=0):
        self._chk(load_library().fall3d_ads_set_tiling(self._h, int(axis), int(cells_per_thread), int(columns),
                                                       int(bins_per_group)), "set_tiling")

    def tiling(self, axis):
        """(cells or bins per thread, bins per workgroup) in use for the sweep of axis."""
        cpt, bpg = _int(), _int()
        self._chk(load_library().fall3d_ads_get_tiling(self._h, int(axis), ctypes.byref(cpt), ctypes.byref(bpg)),
                  "get_tiling")
        return cpt.value, bpg.value

    def autotune(self, dt, stime=0.0, reps=3):
        """Time the candidate tile shapes of each sweep and keep the fastest
        (state restored); returns ((x cells per lane, y bins, z bins), per-sweep ms)."""
        t = (_int * 3)()
        ms = (ctypes.c_double * 3)()
        self._chk(load_library().fall3d_ads_autotune(self._h, float(dt), float(stime), int(reps), t, ms), "autotune")
        return tuple(t), tuple(ms)

    def set_graph(self, on=True):
        self._chk(load_library().fall3d_ads_set_graph(self._h, 1 if on else 0), "set_graph")

    def set_profiling(self, on=True):
        """on: False/0 off, True/1 time every step, N > 1 time one step in N of run()
        (graph replays in between)."""
        self._chk(load_library().fall3d_ads_set_profiling(self._h, int(on)), "set_profiling")

    def kernel_time(self, kernel):
        ms, n = _dbl(), ctypes.c_longlong()
        kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        self._chk(load_library().fall3d_ads_get_kernel_time(self._h, kid, ctypes.byref(ms), ctypes.byref(n)),
                  "get_kernel_time")
        return ms.value, n.value

    # -- learned flux correction ---------------------------------------------------
    def set_flux_output(self, on=True):
        """Accumulate the x/y face mass fluxes of later sweeps (zeroed by this call)."""
        self._chk(load_library().fall3d_ads_set_flux_output(self._h, int(bool(on))), "set_flux_output")

    def face_flux(self, axis):
        """Accumulated face fluxes: axis 0 -> (nx+3, ny, nz, nb), axis 1 -> (nx, ny+3, nz, nb) (x-fastest)."""
        shape = (self.nx + 3, self.ny, self.nz, self.nb) if axis == 0 else (self.nx, self.ny + 3, self.nz, self.nb)
        out = np.zeros(shape, dtype=np.float32, order="F")
        self._chk(load_library().fall3d_ads_get_face_flux(self._h, int(axis), out.ctypes.data), "get_face_flux")
        return out

    def set_learned(self, net=None, smin=0.0, gain=1.0, flux_limit=0.0):
        """Correct the x/y face fluxes of later steps with net (None: plain KT).
        smin < 0: threshold relative to the field maximum; flux_limit > 0: each
        correction bounded by flux_limit times the plain face flux."""
        self._keep["learned"] = net
        self._chk(load_library().fall3d_ads_set_learned(self._h, net._h if net is not None else None, float(smin),
                                                        float(gain), float(flux_limit)), "set_learned")

    def sweep_bytes(self, axis):
        b = _dbl()
        self._chk(load_library().fall3d_ads_sweep_bytes(self._h, int(axis), ctypes.byref(b)), "sweep_bytes")
        return b.value


# ---- module-level singleton, as the reference's private CB (mod_ADS.F90:41) ----
_default: ADS | None = None


class Ensemble:
    """Ensemble members of different sizes advanced together (one batched
    launch per sweep, whole runs replayed from one hipGraph). members[m] is an
    ADS object for member m (fill it with set_problem/set_state, read it with
    get_state/budget); it is owned by the ensemble."""

    def __init__(self, dims, nbins, limiter="minmod", time_marching="rk", real_kind=8, state="fp32"):
        L = load_library()
        dims = [tuple(int(v) for v in d) for d in dims]
        self.nmem, self.nb = len(dims), int(nbins)
        lim = LIMITERS[limiter] if isinstance(limiter, str) else int(limiter)
        tm = TIME_MARCHING[time_marching] if isinstance(time_marching, str) else int(time_marching)
        st = STATES[state] if isinstance(state, str) else int(state)
        flat = (_int * (3 * self.nmem))(*[v for d in dims for v in d])
        h = _vp()
        rc = L.fall3d_ads_ensemble_create(ctypes.byref(h), self.nmem, flat, self.nb, lim, tm, int(real_kind), st)
        self._h = h
        if rc != 0:
            msg = L.fall3d_ads_ensemble_last_error(h).decode()
            L.fall3d_ads_ensemble_release(h)
            self._h = None
            raise Fall3dAdsError(f"fall3d_ads_ensemble_create failed ({rc}): {msg}")
        self.members = []
        for m, d in enumerate(dims):
            mh = _vp()
            self._chk(L.fall3d_ads_ensemble_member(h, m, ctypes.byref(mh)), "ensemble_member")
            self.members.append(ADS(*d, self.nb, lim, tm, real_kind, st, _borrow=mh))

    def _chk(self, rc, what):
        if rc != 0:
            msg = load_library().fall3d_ads_ensemble_last_error(self._h).decode()
            raise Fall3dAdsError(f"{what} failed ({rc}): {msg}")

    def _vec(self, v, default=0.0):
        if v is None:
            v = [default] * self.nmem
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.nmem,)))
        return a, a.ctypes.data_as(_dp)

    def step(self, dt, iiter, stime=None):
        (a, pa), (b, pb) = self._vec(dt), self._vec(stime)
        self._chk(load_library().fall3d_ads_ensemble_step(self._h, pa, pb, int(iiter)), "ensemble_step")

    def run(self, dt, nsteps, iiter0=0, stime=None):
        (a, pa), (b, pb) = self._vec(dt), self._vec(stime)
        self._chk(load_library().fall3d_ads_ensemble_run(self._h, pa, pb, int(iiter0), int(nsteps)), "ensemble_run")

    def synchronize(self):
        self._chk(load_library().fall3d_ads_ensemble_synchronize(self._h), "ensemble_synchronize")

    def set_graph(self, on=True):
        self._chk(load_library().fall3d_ads_ensemble_set_graph(self._h, int(bool(on))), "ensemble_set_graph")

    def set_tiling(self, x_cells_per_lane=0, y_bins_per_group=0, z_bins_per_group=0):
        self._chk(load_library().fall3d_ads_ensemble_set_tiling(self._h, int(x_cells_per_lane), int(y_bins_per_group),
                                                                int(z_bins_per_group)), "ensemble_set_tiling")

    def tiling(self):
        out = (_int * 3)()
        self._chk(load_library().fall3d_ads_ensemble_get_tiling(self._h, out), "ensemble_get_tiling")
        return tuple(out)

    def autotune(self, dt, stime=None, reps=3):
        """Time every candidate tile shape per sweep, keep the fastest; returns
        (tiling, per-sweep milliseconds)."""
        (a, pa), (b, pb) = self._vec(dt), self._vec(stime)
        ms = (ctypes.c_double * 3)()
        self._chk(load_library().fall3d_ads_ensemble_autotune(self._h, pa, pb, int(reps), ms), "ensemble_autotune")
        return self.tiling(), tuple(ms)

    def release(self):
        if getattr(self, "_h", None):
            for mobj in self.members:
                mobj._h = None
            load_library().fall3d_ads_ensemble_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Network:
    """The per-face flux-correction network (BF16 MFMA kernels, FP32 Adam) and
    its training rows on the device."""

    def __init__(self, capacity=0, params=None, seed=0):
        """capacity: training rows kept on the device; params: FP32 parameter
        vector, or None for a He-normal initialisation from seed."""
        L = load_library()
        h = _vp()
        rc = L.fall3d_mlp_create(ctypes.byref(h), int(capacity))
        self._h = h
        if rc != 0:
            msg = L.fall3d_mlp_last_error(h).decode()
            L.fall3d_mlp_release(h)
            self._h = None
            raise Fall3dAdsError(f"fall3d_mlp_create failed ({rc}): {msg}")
        self.nparam = L.fall3d_mlp_nparam()
        self.set_params(self.initial_params(seed) if params is None else params)

    @staticmethod
    def initial_params(seed=0):
        """He-normal weights (std sqrt(2/fan_in); last layer sqrt(1/64)), zero biases,
        in the order W1[32][64], b1[64], W2[64][64], b2[64], w3[64], b3."""
        rng = np.random.default_rng(seed)
        parts = [rng.standard_normal(32 * 64) * np.sqrt(2.0 / 32), np.zeros(64),
                 rng.standard_normal(64 * 64) * np.sqrt(2.0 / 64), np.zeros(64),
                 rng.standard_normal(64) * np.sqrt(1.0 / 64), np.zeros(1)]
        return np.concatenate(parts).astype(np.float32)

    def _chk(self, rc, what):
        if rc != 0:
            msg = load_library().fall3d_mlp_last_error(self._h).decode()
            raise Fall3dAdsError(f"{what} failed ({rc}): {msg}")

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {p.size}")
        self._chk(load_library().fall3d_mlp_set_params(self._h, p.ctypes.data), "mlp_set_params")

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._chk(load_library().fall3d_mlp_get_params(self._h, p.ctypes.data), "mlp_get_params")
        return p

    @staticmethod
    def _rows(x):
        x = np.ascontiguousarray(x, dtype=np.uint16)
        if x.ndim != 2 or x.shape[1] != 32:
            raise ValueError("feature rows must be (n, 32) bfloat16 bit patterns")
        return x

    def forward(self, x):
        x = self._rows(x)
        y = np.zeros(x.shape[0], dtype=np.float32)
        self._chk(load_library().fall3d_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data), "mlp_forward")
        return y

    def gradient(self, x, t):
        x = self._rows(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        sse = ctypes.c_double()
        self._chk(load_library().fall3d_mlp_gradient(self._h, x.ctypes.data, t.ctypes.data, x.shape[0], g.ctypes.data,
                                                     ctypes.byref(sse)), "mlp_gradient")
        return g, sse.value

    def append(self, x, t):
        x = self._rows(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._chk(load_library().fall3d_mlp_data_append(self._h, x.ctypes.data, t.ctypes.data, x.shape[0]),
                  "mlp_data_append")

    def count(self):
        n = ctypes.c_long()
        self._chk(load_library().fall3d_mlp_data_count(self._h, ctypes.byref(n)), "mlp_data_count")
        return n.value

    def rows(self, first=0, n=None):
        n = self.count() - first if n is None else n
        x = np.zeros((n, 32), dtype=np.uint16)
        t = np.zeros(n, dtype=np.float32)
        self._chk(load_library().fall3d_mlp_data_get(self._h, int(first), int(n), x.ctypes.data, t.ctypes.data),
                  "mlp_data_get")
        return x, t

    def refs(self, first=0, n=None):
        """(face refs (n, 4): axis | bin << 1, f, cross index, k; sgn S dxp (n,)) of collected rows."""
        n = self.count() - first if n is None else n
        ref = np.zeros((n, 4), dtype=np.int32)
        den = np.zeros(n, dtype=np.float32)
        self._chk(load_library().fall3d_mlp_data_refs(self._h, int(first), int(n), ref.ctypes.data, den.ctypes.data),
                  "mlp_data_refs")
        return ref, den

    def clear(self):
        self._chk(load_library().fall3d_mlp_data_clear(self._h), "mlp_data_clear")

    def train(self, steps, batch=65536, lr=1e-3, seed=0):
        """steps Adam updates on random batches of the stored rows; returns the
        mean squared error of each batch."""
        loss = np.zeros(max(int(steps), 1), dtype=np.float32)
        self._chk(load_library().fall3d_mlp_train(self._h, int(steps), int(batch), float(lr), int(seed),
                                                  loss.ctypes.data), "mlp_train")
        return loss[:steps]

    def release(self):
        if getattr(self, "_h", None):
            load_library().fall3d_mlp_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def coarse_grain(fine, coarse):
    """Interior of coarse := 2x2 horizontal means of fine."""
    rc = load_library().fall3d_ads_coarse_grain(fine._h, coarse._h)
    coarse._chk(rc, "coarse_grain")


def learned_collect(fine, coarse, net, dt, stime=0.0, iiter=0, smin=0.0):
    """One training sample of the whole coarse grid (see fall3d_ads_learned_collect);
    returns the number of rows added."""
    n = ctypes.c_long()
    rc = load_library().fall3d_ads_learned_collect(fine._h, coarse._h, net._h, float(dt), float(stime), int(iiter),
                                                  float(smin), ctypes.byref(n))
    if rc != 0:
        net._chk(rc, "learned_collect")
    return n.value


def bf16_bits(a):
    """float array -> bfloat16 bit patterns (round to nearest even), uint16."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits):
    """bfloat16 bit patterns -> float32 values."""
    return (np.ascontiguousarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def ADS_initialize(limitertype, time_marching, nx, ny, nz, nbins, grid, real_kind=8, state="fp32"):
    """ADS_initialize with the Domain extents and the grid passed explicitly."""
    global _default
    if _default is not None:
        _default.release()
    _default = ADS(nx, ny, nz, nbins, limitertype, time_marching, real_kind, state, grid=grid)
    return _default


def ADS_get_vcomps():
    return _default.get_vcomps()


def ADS_update_horizontal_vcomps(my_u1, my_u2, my_v1, my_v2, stime):
    _default.update_horizontal_vcomps(my_u1, my_u2, my_v1, my_v2, stime)


def ADS_solve_along_x(my_W_flux, my_E_flux, dt, my_c, my_k1):
    return _default.solve_along_x(my_W_flux, my_E_flux, dt, my_c, my_k1)


def ADS_solve_along_y(my_S_flux, my_N_flux, dt, my_c, my_k2):
    return _default.solve_along_y(my_S_flux, my_N_flux, dt, my_c, my_k2)


def ADS_solve_along_z(my_D_flux, my_U_flux, dt, my_c, my_acum, my_k3, my_w1, my_w2, stime, my_vs):
    return _default.solve_along_z(my_D_flux, my_U_flux, dt, my_c, my_acum, my_k3, my_w1, my_w2, stime, my_vs)


def ADS_release():
    global _default
    if _default is not None:
        _default.release()
        _default = None


def ensemble_statistics(fields, percentiles=(), thresholds=(), logmean=True):
    """Per-point statistics over ensemble members (fall3d_ens_statistics).

    fields: torch tensor (nmem, npoints) of float64 on the GPU (member
    fields on one grid). Returns a dict of GPU tensors: mean, logmean,
    stddev, median, per (len(percentiles), npoints), prob (len(thresholds),
    npoints)."""
    import torch

    if fields.dim() != 2 or fields.dtype != torch.float64 or not fields.is_cuda:
        raise TypeError("fields: a 2-D float64 CUDA tensor (members, points) expected")
    x = fields.contiguous()
    nmem, npts = x.shape
    dev = x.device
    pv = torch.tensor(list(percentiles), dtype=torch.float64, device=dev)
    tv = torch.tensor(list(thresholds), dtype=torch.float64, device=dev)
    out = {k: torch.empty(npts, dtype=torch.float64, device=dev) for k in ("mean", "stddev", "median")}
    out["logmean"] = torch.empty(npts, dtype=torch.float64, device=dev) if logmean else None
    out["per"] = torch.empty((len(percentiles), npts), dtype=torch.float64, device=dev)
    out["prob"] = torch.empty((len(thresholds), npts), dtype=torch.float64, device=dev)

    def ptr(t):
        return None if t is None or t.numel() == 0 else t.data_ptr()

    rc = load_library().fall3d_ens_statistics(
        x.data_ptr(), int(nmem), int(npts), ptr(pv), len(percentiles), ptr(tv), len(thresholds),
        ptr(out["mean"]), ptr(out["logmean"]), ptr(out["stddev"]), ptr(out["median"]), ptr(out["per"]),
        ptr(out["prob"]), torch.cuda.current_stream(dev).cuda_stream)
    if rc != 0:
        raise Fall3dAdsError(f"fall3d_ens_statistics failed ({rc})")
    return out
