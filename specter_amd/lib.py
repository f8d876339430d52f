"""ctypes binding of libspx (include/spx.h).  Loading fails loudly when the
library has not been built -- there is no fallback path."""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import POINTER, c_char_p, c_double, c_int, c_size_t, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "spx.h")

SPX_OK, SPX_ERR_ARG, SPX_ERR_UNSUPPORTED, SPX_ERR_HIP, SPX_ERR_ALLOC, SPX_ERR_STATE = range(6)
SPX_FP64, SPX_FP32_MIXED = 0, 1
SPX_SOLVER_HD, SPX_SOLVER_BOUSS, SPX_SOLVER_ROTBOUSS, SPX_SOLVER_MHD = 0, 1, 2, 3
SPX_BWALL_CONDUCTING, SPX_BWALL_INSULATING = 0, 1


class SpxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libspx error {code}: {msg}")
        self.code = code


class SpxGrid(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("Cz", c_int), ("oz", c_int),
                ("Lx", c_double), ("Ly", c_double), ("Lz", c_double), ("precision", c_int)]


class SpxPlanInfo(ctypes.Structure):
    _fields_ = [("nkx", c_int), ("np", c_int), ("n_spectral", c_size_t), ("n_real", c_size_t),
                ("bytes_spectral", c_size_t), ("bytes_real", c_size_t), ("z_fft_len", c_int),
                ("z_algorithm", c_int), ("device_bytes", c_size_t), ("work_bytes", c_size_t),
                ("work_bytes_bouss", c_size_t), ("state_bytes", c_size_t),
                ("state_bytes_bouss", c_size_t), ("z_threads_per_line", c_int),
                ("z_lines_per_block", c_int), ("z_tables_global", c_int),
                ("z_wave_plan", c_int), ("z_wave_plan_lo", c_int)]


class SpxParams(ctypes.Structure):
    _fields_ = [("solver", c_int), ("ord", c_int), ("dt", c_double), ("nu", c_double),
                ("kappa", c_double), ("xmom", c_double), ("xtemp", c_double),
                ("v_zsta", c_double * 2), ("v_zend", c_double * 2), ("omega", c_double * 3),
                ("mu", c_double), ("b0", c_double * 3), ("bzsta", c_int), ("bzend", c_int)]


_P = c_void_p
_SIGS = {
    "spx_plan_create": [POINTER(SpxGrid), POINTER(c_double), POINTER(c_double), _P, POINTER(_P)],
    "spx_plan_destroy": [_P],
    "spx_plan_get_info": [_P, POINTER(SpxPlanInfo)],
    "spx_plan_sync": [_P],
    "spx_fftp3d_real_to_complex": [_P, _P, _P],
    "spx_fftp3d_complex_to_real": [_P, _P, _P],
    "spx_fftp2d_real_to_complex_xy": [_P, _P, _P],
    "spx_fftp2d_complex_to_real_xy": [_P, _P, _P],
    "spx_fftp1d_real_to_complex_z": [_P, _P],
    "spx_fftp1d_complex_to_real_z": [_P, _P],
    "spx_fc_continue": [_P, _P],
    "spx_derivk": [_P, _P, _P, c_int],
    "spx_laplak": [_P, _P, _P],
    "spx_fc_filter": [_P, _P],
    "spx_gradre": [_P, _P, _P, _P, _P, _P, _P],
    "spx_advect": [_P, _P, _P, _P, _P, _P],
    "spx_gradre_noalloc": [_P, _P, _P, _P, _P, _P, _P],
    "spx_gradre_scratch_bytes": [_P, POINTER(c_size_t), POINTER(c_size_t)],
    "spx_sol_project": [_P, _P, _P, _P, _P, c_int, c_int, c_int],
    "spx_v_imposebc_and_project": [_P, _P, _P, _P, _P, c_int, c_int, POINTER(c_double),
                                   POINTER(c_double)],
    "spx_s_imposebc": [_P, _P],
    "spx_energy": [_P, _P, _P, _P, c_int, POINTER(c_double)],
    "spx_cross": [_P, _P, _P, _P, _P, _P, _P, POINTER(c_double)],
    "spx_helicity": [_P, _P, _P, _P, POINTER(c_double)],
    "spx_divergence": [_P, _P, _P, _P, POINTER(c_double)],
    "spx_bouncheck_z": [_P, _P, _P, POINTER(c_double), POINTER(c_double)],
    "spx_state_create": [_P, POINTER(SpxParams), POINTER(_P)],
    "spx_state_destroy": [_P],
    "spx_state_field": [_P, c_char_p, POINTER(_P)],
    "spx_rkstep1": [_P, _P],
    "spx_substep": [_P, _P, c_int],
    "spx_step": [_P, _P, c_int],
    "spx_state_set_graph": [_P, c_int],
    "spx_variance": [_P, _P, c_int, POINTER(c_double)],
    "spx_product": [_P, _P, _P, POINTER(c_double)],
    "spx_wall_shear": [_P, _P, _P, c_double, POINTER(c_double)],
    "spx_nusselt": [_P, _P, POINTER(c_double)],
    "spx_divergence_max": [_P, _P, _P, _P, POINTER(c_double)],
    "spx_energy_spectrum": [_P, _P, _P, _P, c_int, POINTER(c_double)],
    "spx_plan_create_batch": [POINTER(SpxGrid), POINTER(c_double), POINTER(c_double), c_int, _P,
                              POINTER(_P)],
    "spx_plan_create_fcy": [POINTER(SpxGrid), POINTER(c_double), POINTER(c_double), c_int, c_int,
                            POINTER(c_double), POINTER(c_double), c_int, _P, POINTER(_P)],
    "spx_fftp1d_real_to_complex_x": [_P, _P, _P],
    "spx_fftp2d_real_to_complex_yz": [_P, _P],
    "spx_fftp2d_complex_to_real_yz": [_P, _P],
    "spx_fftp1d_complex_to_real_x": [_P, _P, _P],
    "spx_plan_load_neumann": [_P, c_int, POINTER(c_double), POINTER(c_double)],
    "spx_curlk": [_P, _P, _P, _P, c_int],
    "spx_prodre": [_P, _P, _P, _P, _P, _P, _P],
    "spx_vector": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "spx_a_imposebc_and_project": [_P, _P, _P, _P, _P, c_int, c_int],
    "spx_plan_get_members": [_P, POINTER(c_int), POINTER(c_int)],
    "spx_tune_line": [c_int, c_int, c_int, c_int, c_int],
    "spx_tune_wall": [c_int, c_int, c_int],
    "spx_tune_clear": [],
    "spx_plan_set_active": [_P, c_int],
    "spx_state_set_member": [_P, c_int, c_double, c_double],
    "spx_state_copy_member": [_P, c_int, c_int],
    "spx_prof_enable": [_P, c_int],
    "spx_prof_resolve": [_P, POINTER(c_int)],
    "spx_prof_get": [_P, c_int, c_char_p, c_size_t, POINTER(ctypes.c_long), POINTER(c_double),
                     POINTER(c_double)],
    "spx_dev_alloc": [c_size_t, POINTER(_P)],
    "spx_dev_free": [_P],
    "spx_memcpy_h2d": [_P, _P, _P, c_size_t],
    "spx_memcpy_d2h": [_P, _P, _P, c_size_t],
    "spx_memcpy_d2d": [_P, _P, _P, c_size_t],
    "spx_last_error": [ctypes.c_char_p, c_size_t],
}

_lib = None


def header_symbols() -> list[str]:
    """Entry points declared in include/spx.h."""
    with open(HEADER) as fh:
        return sorted(set(re.findall(r"^int (spx_\w+)\(", fh.read(), flags=re.M)))


def load_library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SpxError(SPX_ERR_STATE,
                       f"{LIB_PATH} is missing: build it with specter_amd.build.build()")
    # torch ships its own HIP runtime and links it under an unversioned name;
    # load it first so libspx binds to the same runtime instance (one device
    # context, shared allocations and streams) instead of pulling in a second.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != SPX_OK:
        buf = ctypes.create_string_buffer(2048)
        load_library().spx_last_error(buf, 2048)
        raise SpxError(rc, buf.value.decode(errors="replace"))
