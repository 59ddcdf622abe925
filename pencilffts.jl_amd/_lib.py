"""ctypes binding of libpencilffts_hip.so (the C ABI of include/pfft.h).

The library is built in-tree by pencilffts.jl_amd/build.py.  There is no
fallback: if it is missing, importing the package fails.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpencilffts_hip.so")

OK, ERR_ARGUMENT, ERR_UNSUPPORTED, ERR_DEVICE, ERR_ALLOC = 0, 1, 2, 3, 4
NOTRANSFORM, FFT, BFFT, RFFT, BRFFT_EVEN, BRFFT_ODD, R2R, INPLACE = 0, 1, 2, 3, 4, 5, 16, 0x100
F32, F64, C32, C64, NOTHING = 0, 1, 2, 3, -1
FORWARD, BACKWARD, BACKWARD_SCALED, ADJOINT_FORWARD, ADJOINT_BACKWARD = 0, 1, 2, 3, 4
FLAG_GENERIC_ONLY, FLAG_NO_AUTOTUNE, FLAG_MFMA_SPLIT16, FLAG_MFMA_BF16X3 = 1, 2, 4, 8
FLAG_REVERSE_TILES = 16
FLAG_STREAM_STORES = 32
FLAG_DENSE_WORK = 64
FLAG_ESTIMATE = 128
NOALLOC_SCRATCH_BYTES = 256 << 20   # static scratch of pfft_execute_noalloc

# FFTW planner flags (FFTW.jl's constants), accepted by PencilFFTPlan
FFTW_MEASURE, FFTW_EXHAUSTIVE, FFTW_PATIENT, FFTW_ESTIMATE, FFTW_WISDOM_ONLY = (
    0, 1 << 3, 1 << 5, 1 << 6, 1 << 21)


def FLAG_CANDIDATE(k):
    """PFFT_FLAG_CANDIDATE(k): run candidate k of every specialised pass."""
    return ((int(k) + 1) & 0xff) << 16

OP_NONE, OP_DERIVATIVE, OP_LAPLACIAN, OP_INV_LAPLACIAN, OP_DEALIAS = 0, 1, 2, 3, 4


class SpectralOp(C.Structure):
    """pfft_spectral_op: a multiplier alpha m(k) fused into pfft_execute_op."""
    _fields_ = [("op", C.c_int32), ("dim", C.c_int32), ("alpha", C.c_double),
                ("accumulate", C.c_int32)]


class BatchItem(C.Structure):
    """pfft_batch_item: one transform of a pfft_execute_batch call."""
    _fields_ = [("plan", C.c_void_p), ("direction", C.c_int32),
                ("op", C.POINTER(SpectralOp)), ("inp", C.c_void_p), ("out", C.c_void_p),
                ("work", C.c_void_p), ("work_bytes", C.c_size_t)]


# every symbol include/pfft.h declares, with (restype, argtypes)
i32, i64, u32, sz, dbl, vp, cp = (C.c_int32, C.c_int64, C.c_uint32, C.c_size_t, C.c_double,
                                  C.c_void_p, C.c_char_p)
P = C.POINTER
SIGNATURES = {
    "pfft_version": (cp, []),
    "pfft_last_error": (cp, []),
    "pfft_transform_check": (C.c_int, [i32]),
    "pfft_transform_length_output": (C.c_int, [i32, i64, P(i64)]),
    "pfft_transform_binv": (C.c_int, [i32, i64, P(i32)]),
    "pfft_transform_scale_factor": (C.c_int, [i32, i64, P(i64)]),
    "pfft_transform_eltype_input": (C.c_int, [i32, i32, P(i32)]),
    "pfft_transform_eltype_output": (C.c_int, [i32, i32, P(i32)]),
    "pfft_transforms_is_inplace": (C.c_int, [i32, P(i32), P(i32)]),
    "pfft_transform_expand_dims": (C.c_int, [i32, i32, P(i32)]),
    "pfft_transform_name": (C.c_int, [i32, C.c_char_p, sz]),
    "pfft_plan_create": (C.c_int, [P(vp), i32, P(i64), P(i32), i32, i32, i32, P(i64), u32]),
    "pfft_plan_destroy": (C.c_int, [vp]),
    "pfft_plan_prepare": (C.c_int, [vp, vp]),
    "pfft_plan_ndims": (C.c_int, [vp, P(i32), P(i32)]),
    "pfft_plan_size_global": (C.c_int, [vp, i32, P(i64)]),
    "pfft_plan_mem_dims": (C.c_int, [vp, i32, P(i64)]),
    "pfft_plan_parent_dims": (C.c_int, [vp, i32, P(i64)]),
    "pfft_plan_permutation": (C.c_int, [vp, i32, P(i32)]),
    "pfft_plan_eltypes": (C.c_int, [vp, P(i32), P(i32)]),
    "pfft_plan_is_inplace": (C.c_int, [vp, P(i32)]),
    "pfft_plan_scale_factor": (C.c_int, [vp, P(dbl)]),
    "pfft_plan_kinds": (C.c_int, [vp, P(i32), P(i32)]),
    "pfft_plan_nbytes": (C.c_int, [vp, P(sz), P(sz), P(sz)]),
    "pfft_plan_stage": (C.c_int, [vp, i32, P(i32), P(i64), P(i32), P(i64), P(i32)]),
    "pfft_plan_describe": (C.c_int, [vp, C.c_char_p, sz]),
    "pfft_execute": (C.c_int, [vp, i32, vp, vp, vp, sz, vp]),
    "pfft_execute_noalloc": (C.c_int, [vp, i32, vp, vp, vp]),
    "pfft_execute_op": (C.c_int, [vp, i32, C.POINTER(SpectralOp), vp, vp, vp, sz, vp]),
    "pfft_execute_many": (C.c_int, [vp, i32, i32, P(vp), P(vp), vp, sz, vp]),
    "pfft_execute_batch": (C.c_int, [i32, P(BatchItem), vp]),
    "pfft_project_divergence_free": (C.c_int, [vp, vp, vp, vp, P(C.c_double), vp]),
    "pfft_tune_export": (C.c_int, [C.c_char_p, sz, P(sz)]),
    "pfft_tune_import": (C.c_int, [C.c_char_p]),
    "pfft_tune_forget": (C.c_int, []),
    "pfft_plan_set_timing": (C.c_int, [vp, i32]),
    "pfft_plan_kernels": (C.c_int, [vp, i32, C.c_char_p, sz]),
    "pfft_plan_stage_times": (C.c_int, [vp, i32, P(C.c_float)]),
}


class PfftError(Exception):
    """A non-zero status of the C ABI."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class ArgumentError(PfftError, ValueError):
    """Raised where PencilFFTs.jl throws ArgumentError."""


class UnsupportedError(PfftError, NotImplementedError):
    pass


class DeviceError(PfftError, RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libpencilffts_hip.so is not built; run `python pencilffts.jl_amd/build.py` "
            "(or __graft_entry__.build()) first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def check(status):
    if status == OK:
        return
    msg = lib.pfft_last_error().decode()
    if status == ERR_ARGUMENT:
        raise ArgumentError(status, msg)
    if status == ERR_UNSUPPORTED:
        raise UnsupportedError(status, msg)
    raise DeviceError(status, msg)
