"""ctypes binding of ``libflorence_amd.so`` (declarations in include/florence_amd.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflorence_amd.so")

FL_MAT_NEOHOOKEAN = 1
FL_MAT_MOONEYRIVLIN = 2
FL_MAT_LINEARELASTIC = 3
FL_MAT_NEARLYINCOMPRESSIBLEMOONEYRIVLIN = 4
FL_MAT_ANISOTROPICMOONEYRIVLIN_1 = 5
FL_S_NONPOSITIVE_J = 1
FL_S_NONFINITE = 2
FL_S_PATTERN = 4
FL_GATHER_BUCKETS = 6
FL_TUNE_STIFF_MINW = 1
FL_TUNE_STIFF_VARIANT_BASE = 2
STIFF_VARIANTS = {0: 1, 1: 4, 2: 3, 3: 5, 4: 3}   # padded node-count class -> number of tile variants
FL_E_ARG = -1
FL_E_HIP = -2
FL_E_SPACE = -3
FL_E_LIMIT = -4
FL_PREC_FP64 = 0
FL_PREC_FP32 = 1
FL_PREC_BF16 = 2
PRECISIONS = {"fp64": FL_PREC_FP64, "fp32": FL_PREC_FP32, "bf16": FL_PREC_BF16}
FL_PREC_CSR32 = 16
FL_TUNE_STIFF_SF = 7
FL_TUNE_GATHER = 8
FL_RECORDS_D = 0
FL_RECORDS_C = 1

# every symbol declared in include/florence_amd.h -> (restype, argtypes)
_vp, _i32, _i64, _sz, _dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t, ctypes.c_void_p
_P64 = ctypes.POINTER(ctypes.c_int64)
_P32 = ctypes.POINTER(ctypes.c_int32)
_PD = ctypes.POINTER(ctypes.c_double)

SIGNATURES = {
    "fl_version": (ctypes.c_int, []),
    "fl_last_error": (ctypes.c_char_p, []),
    "fl_pattern_workspace_size": (_sz, [_i64]),
    "fl_pattern_graph": (ctypes.c_int, [_vp, _i64, _i32, _i64, _vp, _vp, _vp, _P64, _P32, _vp, _sz, _vp]),
    "fl_pattern_fill": (ctypes.c_int, [_vp, _i64, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fl_df_workspace_size": (_sz, [_i32, _i64, _i32]),
    "fl_pattern_gather_plan": (ctypes.c_int, [_vp, _i64, _vp, _P32, _P32, _vp]),
    "fl_df_assemble": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64, _i32,
                                      _vp, _vp, _vp, _vp, _vp, _P32, _P32, _vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    "fl_set_tuning": (ctypes.c_int, [_i32, _i32]),
    "fl_df_workspace_size_ex": (_sz, [_i32, _i64, _i32, _i32]),
    "fl_df_assemble_mt": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64,
                                         _i32, _vp, _vp, _vp, _vp, _vp, _P32, _P32, _vp, _vp, _vp, _i32, _vp, _vp,
                                         _vp, _sz, _vp, _vp]),
    "fl_tensor_basis_bytes": (_sz, [_i32, _i32]),
    "fl_pattern_colour": (ctypes.c_int, [_vp, _i64, _i32, _i64, _i32, _vp, _vp, _vp, _P32]),
    "fl_pattern_scatter_table_bytes": (_sz, [_i64, _i32]),
    "fl_pattern_scatter_table": (ctypes.c_int, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fl_df_workspace_size_cs": (_sz, [_i32, _i64, _i32]),
    "fl_df_assemble_cs": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64,
                                         _i32, _vp, _vp, _vp, _vp, _vp, _P32, _P32, _vp, _vp, _vp, _i32, _vp, _vp,
                                         _i32, _P32, _vp, _vp, _vp, _sz, _vp, _vp]),
    "fl_pattern_rank_table_bytes": (_sz, [_i64, _i32]),
    "fl_pattern_rank_table": (ctypes.c_int, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    "fl_df_assemble_ex": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64, _i32,
                                         _vp, _vp, _vp, _vp, _vp, _P32, _P32, _vp, _vp, _vp, _i32, _vp, _sz, _vp,
                                         _vp]),
    "fl_df_residual_workspace_size": (_sz, [_i32, _i64, _i32]),
    "fl_df_residual": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64,
                                      _vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    "fl_df_residual_tb": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64,
                                         _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    "fl_df_residual_mt": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64,
                                         _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    "fl_df_residual_noalloc": (ctypes.c_int, [_i32, _PD, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32,
                                              _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fl_material_table_stride": (_i32, [_i32]),
    "fl_stiffness_variant_count": (_i32, [_i32]),
    "fl_df_records_form": (_i32, [_i32, _vp, _i32, _i32]),
    "fl_mg_galerkin_workspace_size": (_sz, [_i64, _i32, _i64]),
    "fl_mg_galerkin": (ctypes.c_int, [_vp, _sz, _i32, _i64, _i32, _i32, _vp, _vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp,
                                      _vp, _P32, _P32, _vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    "fl_mg_workspace_size": (_sz, [_vp, _i32]),
    "fl_mf_apply": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "fl_mg_pcg": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _vp, _vp, ctypes.c_double, _i32, _i32, _i32,
                                 ctypes.c_double, _i32, _i32, _i32, _vp, _sz, _P32, ctypes.POINTER(ctypes.c_double),
                                 _vp]),
    "fl_mesh_refine_workspace_size": (_sz, [_i64, _i64, _i32, _i32]),
    "fl_mesh_refine": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _i32, _i32, _vp, _vp, _P64, _vp, _sz, _vp]),
    "fl_csr_spmv": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fl_pcg_workspace_size": (_sz, [_i64]),
    "fl_pcg": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, ctypes.c_double, _i32, _i32, _vp, _sz,
                              _P32, ctypes.POINTER(ctypes.c_double), _vp]),
    "fl_csr_spmv_v32": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fl_pcg_v32": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, ctypes.c_double, _i32, _i32, _vp, _sz,
                                  _P32, ctypes.POINTER(ctypes.c_double), _vp]),
    "fl_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "fl_profile_read": (ctypes.c_int, [_vp, _vp, _i32, _i32]),
    "fl_mlp_param_count": (_i64, []),
    "fl_mlp_workspace_size": (_sz, [_i64]),
    "fl_mlp_forward": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fl_mlp_backward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _sz, _vp]),
    "fl_mlp_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, _i32, _vp]),
    "fl_warmstart_features": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _i64, _vp, _vp, _vp]),
    "fl_warmstart_apply": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp]),
}


def profile_enable(on=True):
    load().fl_profile_enable(1 if on else 0)


def profile_read(max_calls=4096, max_phases=4):
    """[(kind, [phase ms ...]), ...] for the calls recorded since the last read."""
    import numpy as np
    kinds = np.zeros(max_calls, dtype=np.int32)
    ms = np.zeros((max_calls, max_phases), dtype=np.float64)
    n = load().fl_profile_read(kinds.ctypes.data_as(ctypes.c_void_p), ms.ctypes.data_as(ctypes.c_void_p),
                               max_calls, max_phases)
    if n < 0:
        check(n, "fl_profile_read")
    return [(int(kinds[i]), ms[i].tolist()) for i in range(n)]

_lib = None


class FlorenceAMDError(RuntimeError):
    pass


def load():
    """Load the library (once) and attach signatures.  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlorenceAMDError(
            "libflorence_amd.so is not built (expected at {}); run `python -m florence_amd.build`".format(LIB_PATH))
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().fl_last_error().decode(errors="replace")
        raise FlorenceAMDError("{} failed (code {}): {}".format(what, rc, msg))


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def double_array(values):
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return ctypes.cast(arr, _PD), arr
