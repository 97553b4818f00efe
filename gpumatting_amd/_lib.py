"""ctypes binding of libgpumatting.so (the C ABI declared in include/gpumatting.h).

The shared library is built in-tree by `python -m gpumatting_amd.build` (hipcc,
gfx950).  Importing this module does not touch the GPU; the first call that
needs the library loads it and raises GMError if it is missing -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpumatting.so")

GM_OK = 0
STATUS_NAMES = {0: "GM_OK", 1: "GM_ERR_INVALID", 2: "GM_ERR_OOM", 3: "GM_ERR_HIP", 4: "GM_ERR_BREAKDOWN",
                5: "GM_ERR_NONFINITE", 6: "GM_ERR_NOT_CONVERGED", 7: "GM_ERR_UNSUPPORTED"}

METHOD_CG, METHOD_PCG, METHOD_GD = 0, 1, 2
PREC_FP32, PREC_FP16, PREC_BF16 = 0, 1, 2
ASSEMBLY_REFERENCE, ASSEMBLY_ACCURATE = 0, 1


class GMError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class BandedMatrix(C.Structure):
    """Same layout as the reference's BandedMatrix (BandedMatrix.h:23-38)."""
    _fields_ = [("a", C.c_void_p), ("apitch", C.c_size_t), ("bands", C.c_void_p),
                ("nbands", C.c_int), ("rows", C.c_int), ("cols", C.c_int)]


class CompressedMatrix(C.Structure):
    """Same layout as the reference's CompressedMatrix (CompressedMatrix.h:26-37)."""
    _fields_ = [("p", C.c_void_p), ("j", C.c_void_p), ("k", C.c_void_p), ("rows", C.c_int), ("cols", C.c_int),
                ("nnz", C.c_size_t), ("length", C.c_size_t)]


class CoordMatrix(C.Structure):
    """Same layout as the reference's CoordMatrix (CoordMatrix.h:25-36)."""
    _fields_ = [("i", C.c_void_p), ("j", C.c_void_p), ("k", C.c_void_p), ("rows", C.c_int), ("cols", C.c_int),
                ("nnz", C.c_size_t), ("length", C.c_size_t)]


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("restart", C.c_int), ("refine_every", C.c_int), ("check_every", C.c_int), ("verify", C.c_int),
                ("refine_mode", C.c_int)]


class TuneResult(C.Structure):
    _fields_ = [("rows_per_thread", C.c_int), ("tile_rows", C.c_int), ("wg_per_cu", C.c_int),
                ("step_us", C.c_double), ("spmv_rows_per_thread", C.c_int), ("spmv_tile_rows", C.c_int),
                ("spmv_wg_per_cu", C.c_int), ("spmv_us", C.c_double), ("update_wg_per_cu", C.c_int),
                ("update_us", C.c_double), ("spmv_direct", C.c_int), ("step_walk", C.c_int),
                ("fused_us", C.c_double), ("fused_update_us", C.c_double), ("fused_defer_us", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("relres", C.c_double),
                ("true_relres", C.c_double), ("replacements", C.c_int)]


_LIB = None

vp, ci, cf, cd, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
BMp = C.POINTER(BandedMatrix)
CSMp = C.POINTER(CompressedMatrix)
COOp = C.POINTER(CoordMatrix)

_SIGS = {
    "gm_last_error": ([], C.c_char_p),
    "gm_version": ([], C.c_char_p),
    "gm_device_count": ([], ci),
    "gm_bm_copy_to_device": ([BMp, BMp], ci),
    "gm_bm_device_free": ([BMp], ci),
    "gm_bm_device_alloc": ([BMp, ci, ci, vp], ci),
    "gm_bm_copy_to_host": ([BMp, BMp], ci),
    "gm_vec_device_malloc": ([C.POINTER(vp), ci, ci, ci], ci),
    "gm_vec_copy_to_device": ([C.POINTER(vp), vp, ci, ci, ci], ci),
    "gm_vec_device_free": ([vp, ci], ci),
    "gm_bm_ax": ([vp, BMp, vp, vp], ci),
    "gm_bm_axpy": ([vp, BMp, vp, vp, ci, vp], ci),
    "gm_bm_jacobi_sweep": ([vp, vp, BMp, vp, cf, vp], ci),
    "gm_vec_add": ([vp, vp, vp, ci, vp], ci),
    "gm_vec_sub": ([vp, vp, vp, ci, vp], ci),
    "gm_vec_scale_const": ([vp, vp, cf, ci, vp], ci),
    "gm_vec_scale": ([vp, vp, vp, ci, vp], ci),
    "gm_inner_prod": ([vp, vp, vp, ci, vp], ci),
    "gm_inner_prod_noalloc": ([vp, vp, vp, ci, vp], ci),
    "gm_grad_solve": ([vp, BMp, vp, ci, ci, vp], ci),
    "gm_cg_solve": ([vp, BMp, vp, ci, ci, ci, vp], ci),
    "gm_jacobi": ([vp, BMp, vp, cf, ci, ci, vp], ci),
    "gm_gauss_seidel_host": ([vp, BMp, vp, sz], ci),
    "gm_jacobi_host": ([vp, BMp, vp, sz, ci, cf], ci),
    "gm_solve_opts_default": ([C.POINTER(SolveOpts)], None),
    "gm_solve": ([vp, BMp, vp, ci, C.POINTER(SolveOpts), C.POINTER(SolveInfo), vp], ci),
    "gm_true_relres": ([BMp, vp, vp, C.POINTER(cd), vp], ci),
    "gm_levin_laplacian": ([BMp, vp, vp, vp, ci, sz, vp, vp, cf, ci, ci, ci, vp], ci),
    "gm_levin_laplacian_u8": ([BMp, vp, vp, vp, ci, sz, vp, cf, cf, ci, ci, ci, vp], ci),
    "gm_levin_backward": ([vp, vp, vp, vp, ci, sz, vp, vp, vp, vp, cf, ci, ci, vp], ci),
    "gm_host_levin_laplacian": ([BMp, vp, cf, vp, vp, ci, ci, sz], ci),
    "gm_levin_bands": ([ci, vp], None),
    "gm_levin_pad": ([ci], ci),
    "gm_release_workspaces": ([], ci),
    "gm_autotune": ([BMp, ci, ci, C.POINTER(TuneResult), vp], ci),
    "gm_tune_set": ([ci, ci, ci, ci, ci, ci], ci),
    "gm_tune_set_step": ([ci, ci, ci, ci, ci, ci, ci], ci),
    "gm_autotune_report": ([], C.c_char_p),
    "gm_spmv_tune_set": ([ci, ci, ci, ci, ci, ci], ci),
    "gm_batch_create": ([C.POINTER(vp), ci, vp, vp, ci, ci], ci),
    "gm_batch_destroy": ([vp], ci),
    "gm_batch_assemble": ([vp, ci, vp, ci, sz, cf, vp], ci),
    "gm_batch_set_rhs": ([vp, ci, ci, vp, vp, vp, vp], ci),
    "gm_batch_solve": ([vp, C.POINTER(SolveOpts), vp], ci),
    "gm_batch_alpha": ([vp, ci, ci, C.POINTER(vp)], ci),
    "gm_batch_info": ([vp, ci, ci, C.POINTER(SolveInfo)], ci),
    "gm_batch_set_grid": ([vp, ci], ci),
    "gm_batch_set_tiles": ([vp, ci], ci),
    "gm_batch_autotune": ([vp, C.POINTER(SolveOpts), ci, C.POINTER(C.c_double), vp], ci),
    "gm_batch_get_config": ([vp, C.POINTER(ci), C.POINTER(ci), C.POINTER(C.c_double)], ci),
    "gm_batch_tiles": ([vp], ci),
    "gm_vec_copy_to_host": ([vp, vp, ci], ci),
    "gm_free": ([vp], None),
    "gm_pgm_read": ([C.c_char_p, C.POINTER(vp), C.POINTER(ci), C.POINTER(ci), C.POINTER(ci)], ci),
    "gm_ppm_read": ([C.c_char_p, C.POINTER(vp), C.POINTER(ci), C.POINTER(ci), C.POINTER(ci)], ci),
    "gm_pgm_write": ([C.c_char_p, vp, ci, ci, C.c_char_p, ci], ci),
    "gm_quantize_u8": ([vp, ci, vp], None),
    "gm_pgm_write_float": ([C.c_char_p, vp, ci, ci, C.c_char_p, ci], ci),
    "gm_clamped_mse": ([vp, vp, ci, C.POINTER(cd)], ci),
    "gm_bm_dump": ([BMp, C.c_char_p], ci),
    "gm_bm_load": ([BMp, C.c_char_p], ci),
    "gm_csm_init": ([CSMp, ci, ci, sz], ci),
    "gm_csm_free": ([CSMp], ci),
    "gm_coo_init": ([COOp, ci, ci, sz], ci),
    "gm_coo_release": ([COOp], ci),
    "gm_coo_append": ([COOp, vp, vp, vp, sz], ci),
    "gm_coo_to_csm": ([CSMp, COOp, vp], ci),
    "gm_csm_ax": ([vp, CSMp, vp, vp], ci),
    "gm_csm_axpy": ([vp, CSMp, vp, vp, ci, vp], ci),
    "gm_coo_axpy": ([vp, COOp, vp, vp, vp], ci),
    "gm_sp_restriction": ([CSMp, vp, ci, ci, vp], ci),
    "gm_sp_prolong": ([vp, vp, vp, ci, ci, vp], ci),
    "gm_sp_galerkin": ([CSMp, BMp, vp, ci, vp], ci),
    "gm_sp_merge_levels": ([vp, C.POINTER(ci), vp, ci, sz, ci, ci, ci, cf], ci),
    "gm_slic": ([vp, C.POINTER(ci), vp, ci, sz, ci, ci, ci, cf, ci, vp], ci),
    "gm_slic_grid": ([ci, ci, ci, C.POINTER(ci), C.POINTER(ci), C.POINTER(ci)], ci),
    "gm_slic_lab": ([vp, vp, ci, sz, ci, ci, vp], ci),
    "gm_slic_raw": ([vp, vp, ci, vp, ci, ci, ci, cf, ci, vp], ci),
    "gm_slic_connect": ([vp, C.POINTER(ci), ci, ci, ci, vp], ci),
}


def exported_symbols():
    return list(_SIGS)


def lib():
    """Load libgpumatting.so (no device work happens at load)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise GMError(7, f"{LIB_PATH} is not built; run python -m gpumatting_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def check(status: int):
    if status != GM_OK:
        raise GMError(status, lib().gm_last_error().decode())
    return status


def call(name: str, *args):
    return check(getattr(lib(), name)(*args))
