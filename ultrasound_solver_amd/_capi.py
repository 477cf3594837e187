"""ctypes binding of ``libhz.so`` (declarations in ``include/hz.h``).

Loading never falls back to anything: if the library is missing the import
of the binding raises, naming the build step.
"""
import ctypes
from ctypes import POINTER, c_double, c_int32, c_int64, c_void_p, c_char_p
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent/"libhz.so"

HZ_OK, HZ_NOT_CONVERGED, HZ_BAD_ARGUMENT, HZ_OUT_OF_MEMORY, HZ_DEVICE_ERROR = range(5)
FACE_KIND = {"reflecting": 0, "open": 1, "soft": 2, "impedance": 3}
C128, C64 = 0, 1
MODE_A, MODE_RAM, MODE_AH, MODE_RAMH, MODE_RA, MODE_RAH = 0, 1, 2, 3, 4, 5

EXPORTS = ("hz_create", "hz_destroy", "hz_apply", "hz_gmres", "hz_set_sources", "hz_build_rhs",
           "hz_gather_sources", "hz_sample", "hz_device_bytes", "hz_last_error", "hz_set_multigrid",
           "hz_autotune", "hz_set_tile_shapes", "hz_gmres_batch", "hz_gmres_fused", "hz_launch_counts",
           "hz_release_workspace",
           "hz_material_vjp", "hz_basis_field", "hz_basis_adjoint", "hz_basis_field_match", "hz_basis_loss",
           "hz_basis_loss_static", "hz_phase_prologue", "hz_phase_epilogue", "hz_mg_transfer", "hz_mg_level",
           "hz_z_transfer", "hz_surface_create", "hz_surface_destroy", "hz_surface_tables",
           "hz_surface_set_sources", "hz_surface_transform", "hz_surface_solve_flux", "hz_surface_planes",
           "hz_surface_sample", "hz_surface_target_response", "hz_deposit_sources",
           "hz_rasterize_materials", "hz_gaussian_filter")
PRE_JACOBI, PRE_MULTIGRID = 0, 1
LOSS_FIELD_MATCH, LOSS_ENERGY, LOSS_SOFTMAX, LOSS_RATIO = range(4)


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
                ("dx", c_double), ("omega", c_double),
                ("rho", POINTER(c_double)), ("kappa", POINTER(c_double)),
                ("weight", POINTER(c_double)*3), ("stretch", POINTER(c_double)*3),
                ("face_kind", c_int32*6), ("face_term", POINTER(c_double)*6)]


class GmresOpts(ctypes.Structure):
    _fields_ = [("restart", c_int32), ("maxiter", c_int32), ("rtol", c_double),
                ("backward_error_tol", c_double), ("precision", c_int32), ("adjoint", c_int32),
                ("chunk", c_int32), ("max_refinements", c_int32), ("inner_rtol", c_double),
                ("use_graph", c_int32), ("preconditioner", c_int32), ("arnoldi_block", c_int32),
                ("relative_rtol", c_double)]


class LevelDesc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
                ("rho", POINTER(c_double)), ("kappa", POINTER(c_double)),
                ("weight", POINTER(c_double)*3), ("stretch", POINTER(c_double)*3),
                ("edge", POINTER(c_double)*3), ("face_term", POINTER(c_double)*6)]


class Primitive(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("center", c_double*3), ("size", c_double*3)]


class SurfaceDesc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("nz", c_int32), ("dx", c_double),
                ("vx", POINTER(c_double)), ("ix", POINTER(c_double)),
                ("vy", POINTER(c_double)), ("iy", POINTER(c_double)),
                ("shift", POINTER(c_double)), ("z_diagonal", POINTER(c_double)),
                ("z_upper", POINTER(c_double)), ("z_lower", POINTER(c_double))]


class GmresStats(ctypes.Structure):
    _fields_ = [("info", c_int32), ("iterations", c_int32), ("cycles", c_int32),
                ("refinements", c_int32), ("scaled_residual", c_double),
                ("relative_residual", c_double), ("backward_error", c_double),
                ("finite", c_int32), ("converged", c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BasisLossDesc(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("target", c_void_p), ("weight", c_void_p), ("mask", c_void_p),
                ("eps", c_double), ("scale", c_double), ("temperature", c_double),
                ("count_a", c_int64), ("count_b", c_int64)]


class HzError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


_lib = None


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise ImportError(f"{LIBRARY} is missing; build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        handle = ctypes.CDLL(str(LIBRARY))
        handle.hz_create.argtypes = [POINTER(GridDesc), POINTER(c_void_p)]
        handle.hz_destroy.argtypes = [c_void_p]
        handle.hz_apply.argtypes = [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_int64, c_void_p]
        handle.hz_gmres.argtypes = [c_void_p, POINTER(GmresOpts), c_void_p, c_void_p, c_int32, c_int64,
                                    POINTER(GmresStats), c_void_p]
        handle.hz_set_sources.argtypes = [c_void_p, c_int32, c_int64, POINTER(c_int64), POINTER(c_int32),
                                          POINTER(c_double)]
        handle.hz_build_rhs.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_int64, c_void_p]
        handle.hz_gather_sources.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int32, c_void_p]
        handle.hz_sample.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int64, c_void_p,
                                     c_int64, c_int32, c_void_p]
        handle.hz_set_multigrid.argtypes = [c_void_p, c_int32, POINTER(LevelDesc), c_double, c_int32, c_int32,
                                            c_int32, c_int32, c_double, c_int32]
        handle.hz_autotune.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, c_int64, POINTER(c_int32),
                                       POINTER(c_int32), c_int32, c_void_p]
        handle.hz_gmres_batch.argtypes = [c_int32, POINTER(c_void_p), POINTER(GmresOpts), POINTER(c_void_p),
                                          POINTER(c_void_p), POINTER(c_int32), POINTER(c_int64),
                                          POINTER(POINTER(GmresStats)), c_int32, c_void_p]
        handle.hz_gmres_fused.argtypes = [c_int32, POINTER(c_void_p), POINTER(GmresOpts), POINTER(c_void_p),
                                          POINTER(c_void_p), POINTER(c_int32), POINTER(c_int64),
                                          POINTER(POINTER(GmresStats)), c_void_p]
        handle.hz_launch_counts.argtypes = [POINTER(c_int64), c_int32]
        handle.hz_release_workspace.argtypes = [c_void_p]
        handle.hz_basis_field.argtypes = [c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p, c_void_p]
        handle.hz_basis_adjoint.argtypes = [c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p, c_void_p]
        handle.hz_basis_field_match.argtypes = [c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p, c_void_p,
                                                c_double, c_void_p, c_void_p, c_void_p, c_void_p]
        handle.hz_basis_loss.argtypes = [c_void_p, c_int64, c_int64, c_int32, c_void_p, POINTER(BasisLossDesc),
                                         c_void_p, c_void_p, c_void_p, c_void_p]
        handle.hz_basis_loss_static.argtypes = handle.hz_basis_loss.argtypes
        handle.hz_phase_prologue.argtypes = [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]
        handle.hz_phase_epilogue.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_double,
                                             c_double, c_void_p, c_int32, c_void_p, c_void_p]
        handle.hz_mg_transfer.argtypes = [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_void_p]
        handle.hz_mg_level.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, POINTER(c_void_p), POINTER(c_int32),
                                       c_void_p]
        handle.hz_material_vjp.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p,
                                           c_void_p]
        handle.hz_set_tile_shapes.argtypes = [c_void_p, POINTER(c_int32), c_int32]
        handle.hz_rasterize_materials.argtypes = [c_int32, c_int32, c_int32, c_double, c_int32, POINTER(Primitive),
                                                  c_double, c_double, c_double, c_double, c_double, c_void_p,
                                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        handle.hz_gaussian_filter.argtypes = [c_int32, c_int32, c_int32, c_int32, POINTER(c_double), c_void_p,
                                              c_void_p]
        handle.hz_deposit_sources.argtypes = [c_int32, POINTER(c_double), c_int32, POINTER(c_double),
                                              POINTER(c_double), c_double, POINTER(c_int32), c_int32,
                                              POINTER(c_double)]
        handle.hz_z_transfer.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int64, c_void_p,
                                         c_void_p, c_void_p]
        handle.hz_surface_create.argtypes = [POINTER(SurfaceDesc), POINTER(c_void_p)]
        handle.hz_surface_destroy.argtypes = [c_void_p]
        handle.hz_surface_tables.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
        handle.hz_surface_set_sources.argtypes = [c_void_p, c_int32, c_int64, POINTER(c_int64), POINTER(c_int32),
                                                  POINTER(c_double)]
        handle.hz_surface_transform.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p]
        handle.hz_surface_solve_flux.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        handle.hz_surface_planes.argtypes = [c_void_p, POINTER(c_int32), c_int32, c_void_p, c_void_p]
        handle.hz_surface_sample.argtypes = [c_void_p, POINTER(c_int32), POINTER(c_double), c_int32, c_void_p,
                                             c_void_p]
        handle.hz_surface_target_response.argtypes = [c_void_p, POINTER(c_int32), POINTER(c_double), c_int32,
                                                       c_void_p, c_void_p]
        handle.hz_device_bytes.argtypes = [c_void_p]
        handle.hz_device_bytes.restype = c_int64
        handle.hz_last_error.restype = c_char_p
        for name in EXPORTS:
            if name not in ("hz_device_bytes", "hz_last_error"):
                getattr(handle, name).restype = c_int32
        _lib = handle
    return _lib


def check(status, what):
    if status != HZ_OK:
        msg = lib().hz_last_error().decode(errors="replace")
        raise HzError(status, f"{what} failed (status {status}): {msg}")
