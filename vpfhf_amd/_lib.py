"""ctypes declarations of the vfg C-ABI (``include/vfg.h``).

Loading fails loudly when ``libvfg.so`` has not been built: there is no CPU
fallback anywhere in this package.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libvfg.so")

HOST, DEVICE = 0, 1
NONE, ZERO, ONE, FIXED = 0, 1, 2, 3
X0, X1, Y0, Y1, Z0, Z1 = range(6)
ERRORS = {0: "VFG_OK", -1: "VFG_E_ARG", -2: "VFG_E_HIP", -3: "VFG_E_NOMEM", -4: "VFG_E_BREAKDOWN",
          -5: "VFG_E_MAXIT", -6: "VFG_E_NAN", -7: "VFG_E_BUDGET", -8: "VFG_E_NODEV"}
PC_JACOBI, PC_CHEBYSHEV, PC_MG = 0, 1, 2
UNILATERAL_NONE, UNILATERAL_NOCOMPRESSION = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
U_FORM_RECOMPUTE, U_FORM_STORED = 0, 1


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double)]


class Props(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("Gc", C.c_double), ("eta", C.c_double), ("epsilon", C.c_double),
                ("irrevtol", C.c_double), ("atnum", C.c_int), ("hasCrackPressure", C.c_int),
                ("unilateral", C.c_int)]


class BC(C.Structure):
    _fields_ = [("face", C.c_int * 6), ("faceValue", C.c_double * 6),
                ("vertex", C.c_int * 8), ("vertexValue", C.c_double * 8)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("maxit", C.c_int), ("pc", C.c_int),
                ("precision", C.c_int), ("cheb_degree", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("relres", C.c_double),
                ("status", C.c_int)]


class EngineOpts(C.Structure):
    _fields_ = [("max_slots", C.c_int), ("precision", C.c_int), ("u_rtol", C.c_double), ("v_rtol", C.c_double),
                ("maxit", C.c_int), ("graph_chunk", C.c_int), ("use_graph", C.c_int), ("pc", C.c_int)]


class EngineDone(C.Structure):
    _fields_ = [("tag", C.c_int64), ("status", C.c_int), ("am_iterations", C.c_int), ("u_its", C.c_int),
                ("v_outer", C.c_int), ("v_its", C.c_int), ("errV", C.c_double)]


class TunePoint(C.Structure):
    _fields_ = [("schedule", C.c_int), ("oz_scale", C.c_double), ("grid", C.c_int), ("ms", C.c_double)]


class VfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


_lib = None
vp = C.c_void_p
dp = C.POINTER(C.c_double)
H = C.c_void_p

# name -> (restype, argtypes); every function declared in include/vfg.h
PROTOS = {
    "vfg_create": (C.c_int, [C.POINTER(Grid), C.POINTER(Props), C.POINTER(H)]),
    "vfg_destroy": (C.c_int, [H]),
    "vfg_last_error": (C.c_char_p, [H]),
    "vfg_set_props": (C.c_int, [H, C.POINTER(Props)]),
    "vfg_set_bc": (C.c_int, [H, C.POINTER(BC), C.POINTER(BC)]),
    "vfg_set_fields": (C.c_int, [H, vp, vp, vp, vp, vp, vp, C.c_int]),
    "vfg_get_fields": (C.c_int, [H, vp, vp, C.c_int]),
    "vfg_num_nodes": (C.c_int64, [H]),
    "vfg_num_cells": (C.c_int64, [H]),
    "vfg_synchronize": (C.c_int, [H]),
    "vfg_u_apply": (C.c_int, [H, vp, vp, C.c_int]),
    "vfg_u_apply_f32": (C.c_int, [H, vp, vp, C.c_int]),
    "vfg_u_set_form": (C.c_int, [H, C.c_int]),
    "vfg_u_diag": (C.c_int, [H, vp, C.c_int]),
    "vfg_u_pc_apply": (C.c_int, [H, vp, vp, C.c_int]),
    "vfg_u_rhs": (C.c_int, [H, vp, C.c_int]),
    "vfg_u_residual": (C.c_int, [H, vp, vp, vp, C.c_int]),
    "vfg_v_set_u": (C.c_int, [H, vp, C.c_int]),
    "vfg_v_apply": (C.c_int, [H, vp, vp, C.c_int]),
    "vfg_v_diag": (C.c_int, [H, vp, C.c_int]),
    "vfg_v_rhs": (C.c_int, [H, vp, C.c_int]),
    "vfg_v_residual": (C.c_int, [H, vp, vp, C.c_int]),
    "vfg_bc_mask": (C.c_int, [H, C.c_int, vp, C.c_int]),
    "vfg_vec_set_from_bc": (C.c_int, [H, C.c_int, vp, C.c_int]),
    "vfg_vec_apply_dirichlet": (C.c_int, [H, C.c_int, vp, vp, C.c_int]),
    "vfg_energies": (C.c_int, [H, vp, vp, dp, C.c_int]),
    "vfg_energies_static": (C.c_int, [H, vp, vp, dp]),
    "vfg_crack_volume": (C.c_int, [H, vp, vp, dp, vp, C.c_int]),
    "vfg_cell_to_node": (C.c_int, [H, vp, vp, C.c_int, C.c_int]),
    "vfg_u_set_state": (C.c_int, [H, vp, C.c_int]),
    "vfg_energy_fields": (C.c_int, [H, vp, vp, vp, C.c_int]),
    "vfg_penny_build_v": (C.c_int, [H, vp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "vfg_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "vfg_u_solve": (C.c_int, [H, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo), C.c_int]),
    "vfg_v_solve": (C.c_int, [H, vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo), C.c_int]),
    "vfg_batch_create": (C.c_int, [C.POINTER(H), C.c_int, C.POINTER(vp)]),
    "vfg_batch_destroy": (C.c_int, [vp]),
    "vfg_batch_last_error": (C.c_char_p, [vp]),
    "vfg_batch_size": (C.c_int, [vp]),
    "vfg_batch_u_apply": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp)]),
    "vfg_batch_u_apply_f32": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp)]),
    "vfg_batch_time_u_apply": (C.c_int, [vp, C.c_int, C.c_int, dp]),
    "vfg_batch_autotune": (C.c_int, [vp, C.c_int, dp]),
    "vfg_batch_tune_points": (C.c_int, [vp, C.POINTER(TunePoint), C.c_int]),
    "vfg_batch_tuning": (C.c_int, [vp, C.POINTER(TunePoint)]),
    "vfg_batch_set_graph": (C.c_int, [vp, C.c_int, C.c_int]),
    "vfg_batch_u_solve": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(SolveOpts),
                                    C.POINTER(SolveInfo), C.c_int]),
    "vfg_engine_opts_default": (None, [C.POINTER(EngineOpts)]),
    "vfg_engine_create": (C.c_int, [C.POINTER(EngineOpts), C.POINTER(vp)]),
    "vfg_engine_destroy": (C.c_int, [vp]),
    "vfg_engine_last_error": (C.c_char_p, [vp]),
    "vfg_engine_free_slots": (C.c_int, [vp]),
    "vfg_engine_submit": (C.c_int, [vp, H, vp, vp, vp, vp, C.c_int, C.c_double, C.c_int64]),
    "vfg_engine_step": (C.c_int, [vp]),
    "vfg_engine_poll": (C.c_int, [vp, C.POINTER(EngineDone), C.c_int]),
    "vfg_engine_synchronize": (C.c_int, [vp]),
    "vfg_engine_stats": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "vfg_device_bytes": (C.c_int64, [H]),
    "vfg_vec_copy": (C.c_int, [H, vp, vp, C.c_int64]),
    "vfg_vec_pointwise_min": (C.c_int, [H, vp, vp, vp, C.c_int64]),
    "vfg_vec_maxabs_diff": (C.c_int, [H, vp, vp, C.c_int64, dp]),
    "vfg_vec_scale": (C.c_int, [H, vp, C.c_double, C.c_int64]),
    "vfg_vec_set": (C.c_int, [H, vp, C.c_double, C.c_int64]),
    "vfg_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
    "vfg_device_free": (C.c_int, [vp]),
    "vfg_memcpy": (C.c_int, [vp, vp, C.c_size_t, C.c_int]),
    "vfg_time_u_apply": (C.c_int, [H, C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int, C.c_int, dp]),
    "vfg_time_v_apply": (C.c_int, [H, C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int, dp]),
}


def load(path=LIBPATH):
    """Loads libvfg.so; raises if it is missing (build with vpfhf_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libvfg.so not built at %s: run `python -m vpfhf_amd.build`" % path)
    L = C.CDLL(path)
    for name, (res, args) in PROTOS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, h=None):
    if rc != 0:
        msg = load().vfg_last_error(h)
        raise VfgError(rc, msg.decode() if msg else "")
    return rc
