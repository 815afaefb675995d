"""ctypes binding of libe4hip.so (the C ABI declared in include/e4hip.h).

The library is built in-tree by ``eilmer4_amd.build`` (hipcc, gfx950).  There
is no CPU fallback: if the library is missing or cannot be loaded, every entry
point raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libe4hip.so")

E4HIP_OK = 0
ERRORS = {-1: "E4HIP_ERR_ARG", -2: "E4HIP_ERR_HIP", -3: "E4HIP_ERR_HANDLE", -4: "E4HIP_ERR_BAD_CELLS",
          -5: "E4HIP_ERR_DECODE", -6: "E4HIP_ERR_NOT_INIT"}

# every symbol include/e4hip.h declares
EXPORTS = [
    "e4hip_init", "e4hip_finalize", "e4hip_last_error", "e4hip_stream", "e4hip_synchronize",
    "e4hip_gas_ideal", "e4hip_gas_tp", "e4hip_block_create", "e4hip_block_create_geometry", "e4hip_block_destroy",
    "e4hip_block_set_bc",
    "e4hip_block_connect", "e4hip_block_set_axisymmetric", "e4hip_block_set_source_terms",
    "e4hip_block_upload", "e4hip_block_download", "e4hip_block_ncells", "e4hip_config",
    "e4hip_explicit_increment", "e4hip_determine_dt", "e4hip_profile_enable", "e4hip_profile_read",
    "e4hip_gas_tp_set_guess_model", "e4hip_tp_guess", "e4hip_tp_decode_stats", "e4hip_tp_face_stats",
    "e4hip_mlp_grad", "e4hip_mlp_train", "e4hip_block_kchunk", "e4hip_determine_dt_noalloc",
    "e4hip_last_step_bad_cells",
]


class E4hipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class StepConfig(C.Structure):
    """E4hipStepConfig (include/e4hip.h)."""
    _fields_ = [("flux_calculator", C.c_int), ("gasdynamic_update_scheme", C.c_int),
                ("interpolation_order", C.c_int), ("apply_limiter", C.c_int),
                ("extrema_clipping", C.c_int), ("interpolate_in_local_frame", C.c_int),
                ("thermo_interpolator", C.c_int), ("precision_mode", C.c_int),
                ("max_invalid_cells", C.c_int), ("learned_eos_guess", C.c_int), ("batch_blocks", C.c_int),
                ("shear_tolerance", C.c_double), ("compression_tolerance", C.c_double),
                ("stringent_cfl", C.c_int), ("use_graphs", C.c_int), ("M_inf", C.c_double),
                ("autotune", C.c_int)]


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def load(path=LIB_PATH):
    """Load libe4hip.so (no build, no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libe4hip.so is not built (%s); run eilmer4_amd.build.build()" % path)
    L = C.CDLL(path)
    L.e4hip_init.argtypes = [C.c_int]
    L.e4hip_last_error.argtypes = [C.c_char_p, C.c_int]
    L.e4hip_stream.restype = C.c_void_p
    L.e4hip_gas_ideal.argtypes = [C.c_double, C.c_double, _ip]
    L.e4hip_gas_tp.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, _dp, _ip]
    L.e4hip_block_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, _ip]
    L.e4hip_block_create_geometry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip]
    L.e4hip_block_destroy.argtypes = [C.c_int]
    L.e4hip_block_set_bc.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
    L.e4hip_block_kchunk.argtypes = [C.c_int, _ip]
    L.e4hip_block_connect.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.e4hip_block_set_axisymmetric.argtypes = [C.c_int, C.c_double, C.c_double]
    L.e4hip_block_set_source_terms.argtypes = [C.c_int, _dp]
    L.e4hip_block_upload.argtypes = [C.c_int, _dp, _dp]
    L.e4hip_block_download.argtypes = [C.c_int, _dp, _dp]
    L.e4hip_block_ncells.argtypes = [C.c_int, C.POINTER(C.c_long), _ip]
    L.e4hip_config.argtypes = [C.POINTER(StepConfig)]
    L.e4hip_explicit_increment.argtypes = [_ip, C.c_int, C.c_double, _ip]
    L.e4hip_determine_dt.argtypes = [_ip, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
    L.e4hip_determine_dt_noalloc.argtypes = [_ip, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
    L.e4hip_last_step_bad_cells.argtypes = [C.c_int, _ip]
    L.e4hip_profile_enable.argtypes = [C.c_int]
    L.e4hip_profile_read.argtypes = [C.c_int, _dp, C.POINTER(C.c_long)]
    _fp = C.POINTER(C.c_float)
    L.e4hip_gas_tp_set_guess_model.argtypes = [C.c_int, _fp, _fp, _fp, C.c_float, C.c_float]
    L.e4hip_tp_guess.argtypes = [C.c_int, C.c_int, _dp, _dp]
    L.e4hip_tp_decode_stats.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int]
    L.e4hip_tp_face_stats.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int]
    L.e4hip_mlp_grad.argtypes = [_fp, _fp, _fp, C.c_int, _fp, _fp]
    L.e4hip_mlp_train.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _fp]
    _lib = L
    return L


def init(device=0):
    """e4hip_init on first use (idempotent in the library)."""
    check(load().e4hip_init(device))


def last_error():
    buf = C.create_string_buffer(512)
    load().e4hip_last_error(buf, 512)
    return buf.value.decode()


def check(rc):
    if rc != E4HIP_OK:
        raise E4hipError(rc, last_error())
    return rc
