"""Loader of the gfx950 library and ctypes declarations of its C ABI."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "lib", "libsimpois_hip.so")


class SimpoisError(RuntimeError):
    """A status code other than SP_OK returned by the library."""


class sp_params(ctypes.Structure):
    _fields_ = [("tstep", ctypes.c_float), ("niter", ctypes.c_int),
                ("scale", ctypes.c_int), ("cgit", ctypes.c_int),
                ("pmsfac", ctypes.c_double), ("ponlit", ctypes.c_double),
                ("cg_eps", ctypes.c_double), ("trace", ctypes.c_int),
                ("cg_fp32", ctypes.c_int), ("cg_rel", ctypes.c_double),
                ("cg_chunk", ctypes.c_int), ("cg_mg", ctypes.c_int),
                ("autotune", ctypes.c_int), ("guidance", ctypes.c_int),
                ("cg_rel_coarse", ctypes.c_double)]


class sp_image(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int), ("h", ctypes.c_int), ("pd", ctypes.c_int),
                ("g", ctypes.c_void_p), ("f", ctypes.c_void_p),
                ("out", ctypes.c_void_p)]


class sp_stats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("cg_iters_finest", ctypes.c_int),
                ("elapsed_ms", ctypes.c_double)]


class sp_plane_stats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("levels", ctypes.c_int),
                ("holes", ctypes.c_int), ("sweeps", ctypes.c_int),
                ("cg_iters", ctypes.c_int), ("cg_rr", ctypes.c_double),
                ("rel_residual", ctypes.c_double)]


# sp_plane_stats.status
PLANE_SWEPT, PLANE_CONVERGED, PLANE_MAXIT, PLANE_BREAKDOWN, PLANE_NONFINITE = range(5)
PLANE_STATUS_NAMES = ("swept", "converged", "maxit", "breakdown", "nonfinite")


_P = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int

# name -> (restype, argtypes); every symbol declared in include/simpois_hip.h
_SIGNATURES = {
    "poisson_solver_separable": (None, [_P, _P, _P, _I, _I, _I, _F, _I, _I, _F]),
    "poisson_rec": (None, [_P, _P, _P, _I, _I, _F, _I, _I, _I]),
    "poisson_extension_by_cg": (None, [_P, _P, _P, _I, _I, _P, _I, _F]),
    "biharmonic_extension_by_cg": (None, [_P, _P, _P, _I, _I, _P, _I, _F]),
    "main_simpois": (_I, [_I, ctypes.POINTER(ctypes.c_char_p)]),
    "amle": (None, [_P, _P, _I, _I]),
    "main_amle": (_I, [_I, ctypes.POINTER(ctypes.c_char_p)]),
    "linear_diffusion_separable": (None, [_P, _P, _I, _I, _I, _P, _I, _F]),
    "linear_diffusion": (None, [_P, _P, _I, _I, _P, _I, _F]),
    "linear_diffusion_step": (None, [_P, _P, _I, _I, _P, _F]),
    "main_heatd": (_I, [_I, ctypes.POINTER(ctypes.c_char_p)]),
    "iio_read_image_float_vec": (ctypes.POINTER(_F), [ctypes.c_char_p, _P, _P, _P]),
    "iio_read_image_float_split": (ctypes.POINTER(_F), [ctypes.c_char_p, _P, _P, _P]),
    "iio_read_image_float": (ctypes.POINTER(_F), [ctypes.c_char_p, _P, _P]),
    "iio_write_image_float_vec": (None, [ctypes.c_char_p, _P, _I, _I, _I]),
    "iio_write_image_float_split": (None, [ctypes.c_char_p, _P, _I, _I, _I]),
    "iio_write_image_float": (None, [ctypes.c_char_p, _P, _I, _I]),
    "sp_params_default": (None, [ctypes.POINTER(sp_params)]),
    "sp_last_error": (ctypes.c_char_p, []),
    "sp_solve_host": (_I, [_P, _P, _P, _I, _I, _I, ctypes.POINTER(sp_params),
                           ctypes.POINTER(sp_stats)]),
    "sp_solve_device": (_I, [_P, _P, _P, _I, _I, _I, ctypes.POINTER(sp_params),
                             _P, ctypes.POINTER(sp_stats)]),
    "sp_residual_device": (_I, [_P, _P, _P, _I, _I, _I, _I,
                                ctypes.POINTER(ctypes.c_double), _P]),
    "sp_residual_device_noalloc": (_I, [_P, _P, _P, _I, _I, _I, _I,
                                        ctypes.POINTER(ctypes.c_double), _P]),
    "sp_solve_batch_device": (_I, [_I, _P, ctypes.POINTER(sp_params), _P,
                                   ctypes.POINTER(sp_stats)]),
    "sp_solve_batch_device_ex": (_I, [_I, _P, ctypes.POINTER(sp_params), _P,
                                      ctypes.POINTER(sp_stats), _P]),
    "sp_last_cg_trace": (_I, [_I, _I, ctypes.POINTER(_I),
                              ctypes.POINTER(ctypes.c_double), _I]),
    "sp_plan_cache_limit": (None, [_I, ctypes.c_size_t]),
    "sp_plan_cache_clear": (None, []),
    "sp_set_alloc_cap": (ctypes.c_size_t, [ctypes.c_size_t]),
    "sp_device_bytes": (ctypes.c_size_t, []),
    "sp_adjoint_batch_device": (_I, [_I, _P, ctypes.POINTER(sp_params), _P,
                                     ctypes.POINTER(sp_stats)]),
    "sp_profile": (None, [_I]),
    "sp_profile_read": (_I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_long),
                             ctypes.POINTER(ctypes.c_double)]),
    "sp_autotune_report": (_I, [ctypes.c_char_p, _I]),
    "sp_cg_persist_counts": (None, [ctypes.POINTER(ctypes.c_long),
                                    ctypes.POINTER(ctypes.c_long)]),
    "sp_stream_copy": (_I, [_P, _P, ctypes.c_size_t, _P]),
    "sp_plambda_lap_device": (_I, [_P, _P, _I, _I, _I, _I, _I, _P]),
    "sp_plambda_lap_host": (_I, [_P, _P, _I, _I, _I, _I, _I]),
    "sp_amle_device": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P]),
    "sp_amle_host": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I]),
    "sp_heatd_device": (_I, [_P, _P, _P, _I, _I, _I, _I, _F, _P]),
    "sp_heatd_host": (_I, [_P, _P, _P, _I, _I, _I, _I, _F]),
}


def exported_symbols():
    """Names the C ABI must export (the declarations of the public header)."""
    return sorted(_SIGNATURES)


class _Lib:
    _handle = None

    def __getattr__(self, name):
        if _Lib._handle is None:
            # torch ships its own HIP runtime; when torch is present it must
            # be the one the library binds to, so that device pointers of
            # torch tensors and of the library live in one runtime
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            if not os.path.exists(lib_path):
                raise ImportError(
                    f"{lib_path} is missing: build it with "
                    "`make -C imscript_amd/csrc` (or __graft_entry__.build())")
            h = ctypes.CDLL(lib_path)
            for k, (res, args) in _SIGNATURES.items():
                fn = getattr(h, k)
                fn.restype = res
                fn.argtypes = args
            _Lib._handle = h
        return getattr(_Lib._handle, name)


lib = _Lib()


def default_params(**kw):
    """sp_params with the reference defaults (-t 0.25 -n 10 -s 99 -c 0) and
    the PMSFAC / PONLIT / CG_EPS environment knobs, updated by ``kw``."""
    p = sp_params()
    lib.sp_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def check(rc):
    if rc != 0:
        raise SimpoisError(f"simpois_hip status {rc}: "
                           f"{lib.sp_last_error().decode(errors='replace')}")
