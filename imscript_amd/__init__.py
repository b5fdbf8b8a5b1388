"""MI355X (gfx950) implementation of imscript's simpois hole-filling path.

The compute path is the HIP library ``lib/libsimpois_hip.so`` built from
``csrc/``; this package is a thin ctypes layer over its C ABI
(``include/simpois_hip.h``).  Importing it does not require a GPU; calling a
solver does, and fails loudly when the library or the device is missing.

Entry points mirror the reference (``simpois.c:288``, ``:247``,
``cleant_cgpois.c:266,273``):

    poisson_solver_separable(out, inp, dat, w, h, pd, tstep, niter, scale, cgrad)
    poisson_rec(u, g, f, w, h, tstep, niter, scale, cgit)

plus array-level helpers (:func:`simpois`) and device-tensor helpers.
"""
from ._lib import (lib, lib_path, SimpoisError, sp_params, sp_stats,
                   sp_image, sp_plane_stats, default_params, exported_symbols,
                   PLANE_SWEPT, PLANE_CONVERGED, PLANE_MAXIT, PLANE_BREAKDOWN,
                   PLANE_NONFINITE, PLANE_STATUS_NAMES)
from .api import (poisson_solver_separable, poisson_rec,
                  poisson_extension_by_cg, biharmonic_extension_by_cg,
                  simpois, solve_device, solve_batch_device, residual_device,
                  profile, profile_read, autotune_report, stream_copy,
                  plambda_lap, amle, heatd, plan_cache_limit,
                  plan_cache_clear, set_alloc_cap, device_bytes, last_cg_trace,
                  cg_persist_counts)

__all__ = ["lib", "lib_path", "SimpoisError", "sp_params", "sp_stats",
           "sp_image", "default_params", "exported_symbols",
           "poisson_solver_separable", "poisson_rec",
           "poisson_extension_by_cg", "biharmonic_extension_by_cg", "simpois",
           "solve_device", "solve_batch_device", "residual_device", "profile",
           "profile_read", "autotune_report", "stream_copy", "plambda_lap", "amle", "heatd",
           "sp_plane_stats", "PLANE_SWEPT", "PLANE_CONVERGED", "PLANE_MAXIT",
           "PLANE_BREAKDOWN", "PLANE_NONFINITE", "PLANE_STATUS_NAMES",
           "plan_cache_limit", "plan_cache_clear", "set_alloc_cap", "device_bytes",
           "last_cg_trace", "cg_persist_counts"]
