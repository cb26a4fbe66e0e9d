"""ctypes binding of libpower_grid_model_c.so (the MI355X build), same role as the reference's
src/power_grid_model/_core/power_grid_core.py. The library must exist: there is no fallback path."""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("PGM_AMD_LIB", Path(__file__).resolve().parent / "_lib" / "libpower_grid_model_c.so"))

Idx = C.c_int64
ID = C.c_int32
_p = C.c_void_p
_cp = C.c_char_p


def _runtime_files() -> set[str]:
    """Files of the HIP and HSA runtimes mapped into this process."""
    found = set()
    try:
        with open("/proc/self/maps") as maps:
            for line in maps:
                path = line.split()[-1] if line.count(" ") >= 5 else ""
                if "libamdhip64" in path or "libhsa-runtime64" in path:
                    found.add(os.path.realpath(path))
    except OSError:
        pass
    return found


def _share_runtime() -> None:
    """One HIP runtime per process. The library binds libamdhip64.so.7 and libhsa-runtime64.so.1 of the ROCm
    installation. A PyTorch wheel imported afterwards asks for libamdhip64.so and libhsa-runtime64.so by their
    unversioned names, which match no loaded object, and maps its own copies: a second HIP and HSA runtime that
    manage the same GPU beside the library's, each with its own device memory and queues (the tests that use
    torch then ended in GPU memory faults). Loading the installation's runtimes under the unversioned names
    makes those requests resolve to the objects already loaded; when torch was imported first, the library's
    own requests match torch's copies by their versioned names and nothing is loaded twice either."""
    before = _runtime_files()
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        try:
            C.CDLL(name, mode=C.RTLD_GLOBAL)
        except OSError:
            pass
    extra = _runtime_files() - before
    if extra:
        import sys

        print(f"power_grid_model_amd: a second HIP runtime was mapped ({sorted(extra)})", file=sys.stderr)


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python power-grid-model_amd/build.py` "
            "(the power flow runs only through the HIP library)"
        )
    lib = C.CDLL(str(LIB_PATH))
    _share_runtime()
    sig = {
        "PGM_create_handle": (_p, []),
        "PGM_destroy_handle": (None, [_p]),
        "PGM_error_code": (Idx, [_p]),
        "PGM_error_message": (_cp, [_p]),
        "PGM_n_failed_scenarios": (Idx, [_p]),
        "PGM_failed_scenarios": (C.POINTER(Idx), [_p]),
        "PGM_batch_errors": (C.POINTER(_cp), [_p]),
        "PGM_clear_error": (None, [_p]),
        "PGM_version": (_cp, []),
        "PGM_create_options": (_p, [_p]),
        "PGM_destroy_options": (None, [_p]),
        "PGM_set_calculation_type": (None, [_p, _p, Idx]),
        "PGM_set_calculation_method": (None, [_p, _p, Idx]),
        "PGM_set_symmetric": (None, [_p, _p, Idx]),
        "PGM_set_err_tol": (None, [_p, _p, C.c_double]),
        "PGM_set_max_iter": (None, [_p, _p, Idx]),
        "PGM_set_threading": (None, [_p, _p, Idx]),
        "PGM_set_tap_changing_strategy": (None, [_p, _p, Idx]),
        "PGM_set_experimental_features": (None, [_p, _p, Idx]),
        "PGM_amd_set_factor_precision": (None, [_p, _p, Idx]),
        "PGM_meta_n_datasets": (Idx, [_p]),
        "PGM_meta_get_dataset_by_idx": (_p, [_p, Idx]),
        "PGM_meta_dataset_name": (_cp, [_p, _p]),
        "PGM_meta_n_components": (Idx, [_p, _p]),
        "PGM_meta_get_component_by_idx": (_p, [_p, _p, Idx]),
        "PGM_meta_component_name": (_cp, [_p, _p]),
        "PGM_meta_component_size": (C.c_size_t, [_p, _p]),
        "PGM_meta_component_alignment": (C.c_size_t, [_p, _p]),
        "PGM_meta_n_attributes": (Idx, [_p, _p]),
        "PGM_meta_get_attribute_by_idx": (_p, [_p, _p, Idx]),
        "PGM_meta_attribute_name": (_cp, [_p, _p]),
        "PGM_meta_attribute_ctype": (Idx, [_p, _p]),
        "PGM_meta_attribute_offset": (C.c_size_t, [_p, _p]),
        "PGM_is_little_endian": (C.c_int, [_p]),
        "PGM_create_dataset_const": (_p, [_p, _cp, Idx, Idx]),
        "PGM_destroy_dataset_const": (None, [_p]),
        "PGM_dataset_const_add_buffer": (None, [_p, _p, _cp, Idx, Idx, _p, _p]),
        "PGM_dataset_const_set_next_cartesian_product_dimension": (None, [_p, _p, _p]),
        "PGM_dataset_const_add_attribute_buffer": (None, [_p, _p, _cp, _cp, _p]),
        "PGM_create_dataset_mutable": (_p, [_p, _cp, Idx, Idx]),
        "PGM_destroy_dataset_mutable": (None, [_p]),
        "PGM_dataset_mutable_add_buffer": (None, [_p, _p, _cp, Idx, Idx, _p, _p]),
        "PGM_dataset_mutable_add_attribute_buffer": (None, [_p, _p, _cp, _cp, _p]),
        "PGM_create_model": (_p, [_p, C.c_double, _p]),
        "PGM_update_model": (None, [_p, _p, _p]),
        "PGM_copy_model": (_p, [_p, _p]),
        "PGM_get_indexer": (None, [_p, _p, _cp, Idx, _p, _p]),
        "PGM_calculate": (None, [_p, _p, _p, _p, _p]),
        "PGM_destroy_model": (None, [_p]),
        "PGM_amd_batch_create": (_p, [_p, _p, _p, _p]),
        "PGM_amd_batch_run": (None, [_p, _p]),
        "PGM_amd_model_make_resident": (Idx, [_p, _p, _p]),
        "PGM_amd_model_release_device_memory": (None, [_p]),
        "PGM_amd_autotune": (Idx, [_p, _p, _p, _p, _p, _p, Idx]),
        "PGM_amd_set_adjoint_layout": (None, [_p, _p, Idx]),
        "PGM_amd_set_step_refinement": (None, [_p, _p, Idx]),
        "PGM_amd_batch_fetch": (None, [_p, _p, _p]),
        "PGM_amd_batch_set_output": (None, [_p, _p, _p]),
        "PGM_amd_batch_adjoint": (None, [_p, _p, _p, _p, _p, _p]),
        "PGM_amd_batch_destroy": (None, [_p]),
        "PGM_amd_batch_n_scenarios": (Idx, [_p]),
        "PGM_amd_batch_iterations": (None, [_p, _p]),
        "PGM_amd_batch_stats": (None, [_p, _p]),
        "PGM_amd_inspect_topology": (Idx, [_p, _p, _p, _p, _p, Idx]),
        "PGM_amd_plan_stats": (None, [_p, _p, Idx, Idx, Idx, _p]),
        "PGM_amd_minimum_degree_ordering": (Idx, [Idx, _p, _p, _p, _p, _p, _p, Idx]),
        "PGM_amd_block_lu_solve": (C.c_int, [Idx, Idx, _p, _p, _p, Idx, _p, _p, _p, _p, Idx]),
        "PGM_amd_block_lu_solve_noalloc": (C.c_int, [Idx, Idx, _p, _p, _p, Idx, _p, _p, _p, _p, Idx]),
        "PGM_amd_copy_probe": (C.c_int, [Idx]),
        "PGM_amd_tap_regulator_ranks": (None, [_p, _p, _p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


class Handle:
    """Owns a PGM_Handle; raises the error it holds after each call (power_grid_core.py error handling)."""

    def __init__(self):
        self.ptr = lib.PGM_create_handle()

    def __del__(self):
        if getattr(self, "ptr", None):
            lib.PGM_destroy_handle(self.ptr)
            self.ptr = None

    def code(self) -> int:
        return lib.PGM_error_code(self.ptr)

    def message(self) -> str:
        return lib.PGM_error_message(self.ptr).decode()

    def check(self):
        from .errors import raise_regular_error

        if self.code() != 0:
            raise_regular_error(self.message())


_handle = None


def handle() -> Handle:
    global _handle
    if _handle is None:
        _handle = Handle()
    return _handle
