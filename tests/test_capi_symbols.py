"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import re

from conftest import ROOT


def declared_functions():
    text = (ROOT/"include"/"hz.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:hz_status|int64_t|const char \*)\s*(hz_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert {"hz_create", "hz_destroy", "hz_apply", "hz_gmres", "hz_set_sources", "hz_build_rhs",
            "hz_gather_sources", "hz_sample", "hz_device_bytes", "hz_last_error"} <= set(names)


def test_library_exports_every_declared_symbol():
    from ultrasound_solver_amd import _capi
    lib = _capi.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(_capi.EXPORTS) == set(declared_functions())


def test_struct_layouts_match_header():
    from ultrasound_solver_amd import _capi
    # int32 x3, pad, 2 doubles, 2 ptrs, 3+3 ptrs, int32 x6, 6 ptrs
    assert ctypes.sizeof(_capi.GridDesc) == 16 + 16 + 16 + 48 + 24 + 48
    assert ctypes.sizeof(_capi.GmresOpts) == 4*2 + 8*2 + 4*4 + 8 + 4*3 + 4 + 8  # 4 bytes of padding before relative_rtol
    assert ctypes.sizeof(_capi.LevelDesc) == 16 + 16 + 24*3 + 48
    assert ctypes.sizeof(_capi.GmresStats) == 4*4 + 8*3 + 4*2
    assert ctypes.sizeof(_capi.SurfaceDesc) == 4*3 + 4 + 8 + 8*8  # 4 bytes of padding before dx
    assert ctypes.sizeof(_capi.Primitive) == 4 + 4 + 8*6  # 4 bytes of padding before center
    assert ctypes.sizeof(_capi.BasisLossDesc) == 4 + 4 + 8*3 + 8*3 + 8*2  # 4 bytes of padding after kind


def test_last_error_is_empty_string_initially():
    from ultrasound_solver_amd import _capi
    assert isinstance(_capi.lib().hz_last_error(), bytes)


def test_bad_arguments_are_rejected_without_touching_the_device():
    from ultrasound_solver_amd import _capi
    handle = ctypes.c_void_p()
    status = _capi.lib().hz_create(None, ctypes.byref(handle))
    assert status == _capi.HZ_BAD_ARGUMENT
    assert b"null" in _capi.lib().hz_last_error()
    # The fused multi-scene solve validates its arrays before any device call.
    status = _capi.lib().hz_gmres_fused(0, None, None, None, None, None, None, None, None)
    assert status == _capi.HZ_BAD_ARGUMENT
    assert b"nscenes" in _capi.lib().hz_last_error()
    # The optimiser-step kernels check pointers, sizes and the target count.
    lib = _capi.lib()
    assert lib.hz_phase_prologue(None, 4, None, 0, None, None, None) == _capi.HZ_BAD_ARGUMENT
    fake = ctypes.c_void_p(16)
    assert lib.hz_phase_prologue(fake, 513, None, 0, fake, None, None) == _capi.HZ_BAD_ARGUMENT
    assert b"512" in lib.hz_last_error()
    assert lib.hz_phase_prologue(fake, 4, None, 2, fake, None, None) == _capi.HZ_BAD_ARGUMENT
    assert b"targets" in lib.hz_last_error()
    assert lib.hz_phase_epilogue(fake, 4, None, None, 0, None, 1.0, 1e-12, fake, 1, fake, None) == \
        _capi.HZ_BAD_ARGUMENT
    assert b"gradient" in lib.hz_last_error()
    assert lib.hz_phase_epilogue(fake, 4, fake, None, 0, None, 1.0, 0.0, fake, 1, fake, None) == \
        _capi.HZ_BAD_ARGUMENT
    assert b"eps" in lib.hz_last_error()


def test_operator_exposes_every_entry_point():
    """Each C entry point has its Python method (no compute: attribute check)."""
    from ultrasound_solver_amd import operator
    for name in ("apply", "gmres", "autotune", "set_tile_shapes", "material_vjp", "release_workspace",
                 "set_multigrid", "set_sources", "build_rhs", "gather_sources", "sample", "device_bytes"):
        assert callable(getattr(operator.HelmholtzOperator, name)), name
    assert callable(operator.gmres_batch)
    assert callable(operator.gmres_fused)
    assert callable(operator.launch_counts)
