"""ctypes binding of ``libbosss_hip.so`` (the C ABI declared in include/bosss_hip.h).

The product path has no CPU fallback: if the HIP library is missing this module
raises on import of the operator API.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbosss_hip.so")

BOSSS_FORM_SIP, BOSSS_FORM_ADVDIFF, BOSSS_FORM_EULER_LF, BOSSS_FORM_EULER_HLLC = 0, 1, 2, 3
BOSSS_FP64, BOSSS_FP32 = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class GridDesc(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int32), ("J", ctypes.c_int32), ("E", ctypes.c_int32),
                ("detJ", _dp), ("invJ", _dp), ("scaling", _dp), ("lenscale", _dp),
                ("cells", _ip), ("trafo", _ip), ("normals", _dp), ("sqrtgram", _dp)]


class TablesDesc(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int32), ("degree", ctypes.c_int32), ("N", ctypes.c_int32), ("Kv", ctypes.c_int32),
                ("Ke", ctypes.c_int32), ("ntrafo", ctypes.c_int32),
                ("vol_weights", _dp), ("phi", _dp), ("dphi", _dp), ("face_weights", _dp), ("trace", _dp), ("dtrace", _dp)]


class FormParams(ctypes.Structure):
    _fields_ = [("penalty_base", ctypes.c_double), ("nu", ctypes.c_double), ("velocity", ctypes.c_double * 3),
                ("gamma", ctypes.c_double), ("mach", ctypes.c_double)]


_i64p = ctypes.POINTER(ctypes.c_int64)


class XdgDesc(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int32), ("degree", ctypes.c_int32), ("N", ctypes.c_int32), ("J", ctypes.c_int32),
                ("mode_exponents", _ip), ("nspec", _ip), ("spec0", _ip), ("invJ", _dp), ("scaling", _dp), ("lscale", _dp),
                ("penalty_base", ctypes.c_double), ("mu", ctypes.c_double * 2), ("source", ctypes.c_double * 2),
                ("S", ctypes.c_int32), ("cols", _ip), ("boff", _i64p),
                ("nvol", ctypes.c_int32), ("vol_cell", _ip), ("vol_spec", _ip), ("vol_off", _i64p), ("vol_nodes", _dp), ("vol_w", _dp),
                ("nface", ctypes.c_int32), ("face_cells", _ip), ("face_slot", _ip), ("face_spec", _ip), ("face_off", _i64p),
                ("face_nodes_in", _dp), ("face_nodes_out", _dp), ("face_w", _dp), ("face_normal", _dp), ("face_gD", _dp),
                ("nifc", ctypes.c_int32), ("ifc_cell", _ip), ("ifc_off", _i64p), ("ifc_nodes", _dp), ("ifc_w", _dp),
                ("ifc_normal", _dp)]


# symbol -> (restype, argtypes); must match include/bosss_hip.h
SIGNATURES = {
    "bosss_hip_last_error": (ctypes.c_char_p, []),
    "bosss_hip_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "bosss_hip_ctx_destroy": (ctypes.c_int, [_vp]),
    "bosss_hip_sync": (ctypes.c_int, [_vp]),
    "bosss_hip_ctx_stream": (_vp, [_vp]),
    "bosss_hip_grid_upload": (ctypes.c_int, [_vp, ctypes.POINTER(GridDesc), ctypes.POINTER(_vp)]),
    "bosss_hip_grid_destroy": (ctypes.c_int, [_vp]),
    "bosss_hip_op_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(TablesDesc), ctypes.c_int, ctypes.POINTER(FormParams),
                                            ctypes.c_int, ctypes.POINTER(_vp)]),
    "bosss_hip_op_destroy": (ctypes.c_int, [_vp]),
    "bosss_hip_op_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32)]),
    "bosss_hip_op_kernel_family": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "bosss_hip_evaluate": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]),
    "bosss_hip_jvp": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "bosss_hip_vjp": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "bosss_hip_evaluate_dev": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]),
    "bosss_hip_jvp_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_vjp_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_evaluate_vjp_dev": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_op_matrix_layout": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                                  ctypes.POINTER(ctypes.c_int32), _vp]),
    "bosss_hip_op_assemble_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "bosss_hip_op_diag_blocks_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _vp, _vp]),
    "bosss_hip_bell_spmv_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp]),
    "bosss_hip_op_set_spmv_tile": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "bosss_hip_op_spmv_tiles": (ctypes.c_int, [_vp, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_float), ctypes.c_int32]),
    "bosss_hip_block_lu_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp]),
    "bosss_hip_block_lu_noalloc_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32,
                                                      _vp]),
    "bosss_hip_block_solve_dev": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_block_gemv_dev": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp]),
    "bosss_hip_block_last_error": (ctypes.c_char_p, []),
    "bosss_hip_xdg_create": (ctypes.c_int, [_vp, ctypes.POINTER(XdgDesc), ctypes.POINTER(_vp)]),
    "bosss_hip_xdg_destroy": (ctypes.c_int, [_vp]),
    "bosss_hip_xdg_info": (ctypes.c_int, [_vp, _i64p, _i64p]),
    "bosss_hip_xdg_assemble_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "bosss_hip_xdg_spmv_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp]),
    "bosss_hip_xdg_solve_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                               _dp, _vp]),
    "bosss_hip_xdg_set_tile": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "bosss_hip_xdg_tiles": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)]),
    "bosss_hip_xdg_last_error": (ctypes.c_char_p, []),
    "bosss_hip_measure_peaks": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double)]),
    "bosss_hip_measure_fp64_mix": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "bosss_hip_peaks_last_error": (ctypes.c_char_p, []),
    "bosss_hip_gmres_workspace": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "bosss_hip_gmres_arnoldi_dev": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                                                   ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_gmres_update_dev": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                                                  ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_block_gemv_uniform_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_double, _vp,
                                                        ctypes.c_double, _vp, _vp]),
    "bosss_hip_krylov_last_error": (ctypes.c_char_p, []),
    "bosss_hip_buffer_create": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "bosss_hip_buffer_destroy": (ctypes.c_int, [_vp, _vp]),
    "bosss_hip_buffer_upload": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64]),
    "bosss_hip_buffer_download": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64]),
    "bosss_hip_op_set_check_finite": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "bosss_hip_op_assemble": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, _vp]),
    "bosss_hip_bell_spmv": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp]),
    "bosss_hip_block_lu": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32]),
    "bosss_hip_block_solve": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bosss_hip_block_gemv": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp]),
    "bosss_hip_xdg_assemble": (ctypes.c_int, [_vp, _vp, _vp]),
    "bosss_hip_xdg_spmv": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp]),
    "bosss_hip_xdg_solve": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _dp]),
}

_lib = None


class BosssError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bosss_hip error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("libbosss_hip.so not found at %s; run __graft_entry__.build() first" % LIBPATH)
        # PyTorch-ROCm ships its own HIP runtime (DT_NEEDED "libamdhip64.so", soname
        # libamdhip64.so.7).  Loading torch first lets this library bind to that same runtime
        # by soname; loading this library first would put a second HIP runtime into the
        # process, after which torch finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().bosss_hip_last_error().decode()
        raise BosssError(rc, msg)
    return rc


def device_array(t, dtype, numel, device, name, at_least=False):
    """Refuse a tensor a kernel would read or write out of bounds: it must be a contiguous
    CUDA tensor of `dtype` on device index `device` with exactly (or, with at_least, at
    least) `numel` entries.  The C entry points receive bare pointers and cannot check."""
    ok_size = t.numel() >= numel if at_least else t.numel() == numel
    if not (t.is_cuda and t.dtype == dtype and t.is_contiguous() and ok_size and t.device.index == device):
        raise ValueError("%s: expected a contiguous %s tensor on cuda:%d with %s%d entries, got %s %s on %s%s"
                         % (name, dtype, device, "at least " if at_least else "", numel, t.dtype, tuple(t.shape), t.device,
                            "" if t.is_contiguous() else " (not contiguous)"))
    return t
