"""ctypes binding of libampsolve.so (include/ampsolve.h). Loading fails loudly when the library is missing."""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libampsolve.so")

AMP_OK = 0
AMP_ERR_ARG = -1
AMP_ERR_HIP = -2
AMP_ERR_STATE = -3
AMP_ERR_PATTERN = -4
AMP_ERR_INTERNAL = -5
AMP_CHOLESKY = 0
AMP_LDLT = 1
AMP_F64 = 0
AMP_F32 = 1
AMP_F32_BF16X3 = 2

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class SymInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("nsuper", C.c_int32), ("maxFront", C.c_int32), ("nlevels", C.c_int32),
                ("nnzL", C.c_int64), ("nnzA", C.c_int64), ("nnzLStored", C.c_int64), ("nnzPattern", C.c_int64),
                ("flops", C.c_double), ("fill", C.c_double), ("gemmFlops", C.c_double),
                ("bulkBytes", C.c_double), ("bulkLaunches", C.c_int32), ("patternKey", C.c_char * 64)]


class SymArrays(C.Structure):
    _fields_ = [(name, C.POINTER(C.c_int32)) for name in
                ("perm", "pinv", "parent", "colCount", "lowerP", "lowerI", "lowerSrc", "superStart", "colSuper",
                 "superParent", "childHead", "childNext", "Sp", "Si")] + [("Lxp", C.POINTER(C.c_int64))]


class SymDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("nsuper", C.c_int32), ("nnzPattern", C.c_int64)] + \
               [(name, C.c_void_p) for name in ("perm", "parent", "colCount", "superStart", "Sp", "Si", "lowerP",
                                                 "lowerI", "lowerSrc")] + [("patternKey", C.c_char_p)]


class RefineOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("maxSteps", C.c_int32), ("restart", C.c_int32), ("tol", C.c_double)]


class RefineInfo(C.Structure):
    _fields_ = [("steps", C.c_int32), ("converged", C.c_int32), ("relResidual", C.c_double), ("ms", C.c_double)]


AMP_REFINE_IR = 0
AMP_REFINE_GMRES = 1


class FactorInfo(C.Structure):
    _fields_ = [("ok", C.c_int32), ("nNeg", C.c_int32), ("nPerturbed", C.c_int32), ("failColumn", C.c_int32),
                ("ms", C.c_double)]


class TuneInfo(C.Structure):
    _fields_ = [("smallMax", C.c_int32), ("smallNarrow", C.c_int32), ("wideLaunch", C.c_int32),
                ("superPanel", C.c_int32), ("flowWideMin", C.c_int32), ("packMin", C.c_int32),
                ("fusedPanel", C.c_int32), ("fusedSolve", C.c_int32),
                ("candidates", C.c_int32),
                ("defaultMs", C.c_double), ("bestMs", C.c_double)]


class PcgOpts(C.Structure):
    _fields_ = [("maxIter", C.c_int32), ("tol", C.c_double)]


class PcgInfo(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("relResidual", C.c_double),
                ("initialResidual", C.c_double), ("ms", C.c_double)]


f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")

# every entry point of include/ampsolve.h with its ctypes signature
SIGNATURES = {
    "amp_amd_order": ([C.c_int32, i32p, i32p, C.c_double, i32p], C.c_int),
    "amp_nd_order": ([C.c_int32, i32p, i32p, f64p, f64p, C.c_int32, i32p], C.c_int),
    "amp_analyze": ([C.c_int32, i32p, i32p, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "amp_sym_import": ([C.POINTER(SymDesc), C.POINTER(C.c_void_p)], C.c_int),
    "amp_pattern_fingerprint": ([C.c_int32, i32p, i32p, C.c_char_p], C.c_int),
    "amp_sym_get_info": ([C.c_void_p, C.POINTER(SymInfo)], C.c_int),
    "amp_sym_get_arrays": ([C.c_void_p, C.POINTER(SymArrays)], C.c_int),
    "amp_sym_free": ([C.c_void_p], None),
    "amp_num_create": ([C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "amp_factor": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.POINTER(FactorInfo)], C.c_int),
    "amp_factor_device": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.POINTER(FactorInfo)], C.c_int),
    "amp_solve": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_solve_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_solve_device_noalloc": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_refine": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(RefineOpts), C.POINTER(RefineInfo)],
                   C.c_int),
    "amp_refine_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(RefineOpts),
                           C.POINTER(RefineInfo)], C.c_int),
    "amp_sym_set_pattern": ([C.c_void_p, i32p, i32p], C.c_int),
    "amp_download_factor": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "amp_sym_autotune": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(TuneInfo)], C.c_int),
    "amp_sym_set_params": ([C.c_void_p, C.POINTER(TuneInfo)], C.c_int),
    "amp_shift_invert_apply": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_shift_invert_apply_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_pcg_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PcgOpts), C.POINTER(PcgInfo)],
                       C.c_int),
    "amp_pcg_batch_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                              C.POINTER(PcgOpts), C.POINTER(PcgInfo)], C.c_int),
    "amp_mlp_create": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
    "amp_mlp_param_count": ([C.c_void_p], C.c_int64),
    "amp_mlp_set_params": ([C.c_void_p, f32p, C.c_int64], C.c_int),
    "amp_mlp_get_params": ([C.c_void_p, f32p, C.c_int64], C.c_int),
    "amp_mlp_get_grads": ([C.c_void_p, f32p, C.c_int64], C.c_int),
    "amp_mlp_forward": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "amp_mlp_gradient": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "amp_mlp_adam": ([C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
    "amp_mlp_free": ([C.c_void_p], None),
    "amp_num_release_workspace": ([C.c_void_p], C.c_int),
    "amp_graph_create": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                          C.POINTER(C.c_void_p)], C.c_int),
    "amp_graph_launch": ([C.c_void_p], C.c_int),
    "amp_graph_sync": ([C.c_void_p, C.POINTER(FactorInfo)], C.c_int),
    "amp_graph_free": ([C.c_void_p], None),
    "amp_num_set_profiling": ([C.c_void_p, C.c_int32], C.c_int),
    "amp_num_get_kernel_ms": ([C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)], C.c_int),
    "amp_num_set_growth_limit": ([C.c_void_p, C.c_double], C.c_int),
    "amp_num_free": ([C.c_void_p], None),
    "amp_asm_create": ([C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)],
                       C.c_int),
    "amp_asm_get_pattern": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.POINTER(C.c_int32)),
                            C.POINTER(C.POINTER(C.c_int32))], C.c_int),
    "amp_assemble_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_asm_set_modes": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "amp_param_coefficients_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "amp_asm_free": ([C.c_void_p], None),
    "amp_last_error": ([], C.c_char_p),
}

_lib = None


class AmpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """The loaded library; raises when libampsolve.so has not been built (no silent fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libampsolve.so not built at {LIB_PATH}: run `make -C applied-multiphysics_amd` "
                               "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != AMP_OK:
        raise AmpError(rc, lib().amp_last_error().decode())
    return rc


def ptr(a):
    """Address of a numpy array, a device array (ampsolve.device.DeviceArray) or a framework device tensor. The
    library's device entry points order their work after the null stream; for a framework tensor whose producer may
    still be queued on another current stream of that framework, that stream is synchronised here."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if getattr(a, "is_cuda", False) and type(a).__module__.startswith("torch"):
            import torch

            s = torch.cuda.current_stream(a.device)
            if s.cuda_stream != 0:
                s.synchronize()
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)
