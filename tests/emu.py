"""ctypes wrapper of libpencilffts_emu.so: the host emulator of the generic
pass (the kernel's phase functions run thread by thread on the CPU).  Used by
the CPU tests to check the kernel arithmetic and index maps against the
oracle without a GPU."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.environ.get("PFFT_EMU_LIB") or os.path.join(ROOT, "pencilffts.jl_amd", "lib", "libpencilffts_emu.so")

NP_DTYPES = {0: np.float32, 1: np.float64, 2: np.complex64, 3: np.complex128}


class Emulator:
    def __init__(self):
        if not os.path.exists(PATH):
            raise RuntimeError("emulator library not built (run __graft_entry__.build())")
        self.lib = C.CDLL(PATH)
        L = self.lib
        L.pfft_emu_plan_create.restype = C.c_int
        L.pfft_emu_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32,
                                           C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                           C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(C.c_int64), C.c_uint32]
        L.pfft_emu_plan_destroy.argtypes = [C.c_void_p]
        L.pfft_emu_last_error.restype = C.c_char_p
        L.pfft_emu_nbytes.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.pfft_emu_execute.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.pfft_emu_set_traversal.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.pfft_emu_execute_op.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                          C.c_void_p]

    def run(self, dims, kinds, real_type, direction, x, permute=True, extra=(), traversal=(0, 0),
            flags=0):
        """x: numpy array holding the column-major memory image of the input
        of `direction`; returns the output memory image (flat).  traversal:
        (tile order, descending) of every pass; flags: plan flags."""
        N = len(dims)
        h = C.c_void_p()
        ne = len(extra)
        st = self.lib.pfft_emu_plan_create(
            C.byref(h), N, (C.c_int64 * N)(*dims), (C.c_int32 * N)(*kinds), real_type,
            int(permute), ne, (C.c_int64 * max(ne, 1))(*extra), flags)
        if st:
            raise RuntimeError(self.lib.pfft_emu_last_error().decode())
        try:
            self.lib.pfft_emu_set_traversal(h, *traversal)
            bi, bo = C.c_size_t(), C.c_size_t()
            self.lib.pfft_emu_nbytes(h, C.byref(bi), C.byref(bo))
            nin, nout = (bi.value, bo.value) if direction in (0, 4) else (bo.value, bi.value)
            src = np.ascontiguousarray(x).reshape(-1)
            assert src.nbytes == nin, (src.nbytes, nin)
            out = np.zeros(nout, dtype=np.uint8)
            st = self.lib.pfft_emu_execute(h, direction, src.ctypes.data, out.ctypes.data)
            if st:
                raise RuntimeError(self.lib.pfft_emu_last_error().decode())
            return out
        finally:
            self.lib.pfft_emu_plan_destroy(h)

    def bluestein(self, dims, kinds, real_type, permute=True):
        """Convolution length of the Bluestein form of each forward pass (0:
        direct radix stages)."""
        N = len(dims)
        h = C.c_void_p()
        st = self.lib.pfft_emu_plan_create(
            C.byref(h), N, (C.c_int64 * N)(*dims), (C.c_int32 * N)(*kinds), real_type,
            int(permute), 0, (C.c_int64 * 1)(), 0)
        if st:
            raise RuntimeError(self.lib.pfft_emu_last_error().decode())
        try:
            out = (C.c_int64 * 16)()
            self.lib.pfft_emu_bluestein.restype = C.c_int
            n = self.lib.pfft_emu_bluestein(h, out, 16)
            return [out[i] for i in range(n)]
        finally:
            self.lib.pfft_emu_plan_destroy(h)

    def run_inplace(self, dims, kinds, real_type, directions, buf, permute=True, extra=()):
        """In-place plan: runs each direction in turn on the uint8 buffer `buf`
        (input and output views share it); returns the plan's buffer size."""
        N = len(dims)
        h = C.c_void_p()
        ne = len(extra)
        st = self.lib.pfft_emu_plan_create(
            C.byref(h), N, (C.c_int64 * N)(*dims), (C.c_int32 * N)(*kinds), real_type,
            int(permute), ne, (C.c_int64 * max(ne, 1))(*extra), 0)
        if st:
            raise RuntimeError(self.lib.pfft_emu_last_error().decode())
        try:
            bi, bo = C.c_size_t(), C.c_size_t()
            self.lib.pfft_emu_nbytes(h, C.byref(bi), C.byref(bo))
            assert bi.value == bo.value and buf.nbytes >= bi.value, (bi.value, bo.value)
            for d in directions:
                st = self.lib.pfft_emu_execute(h, d, buf.ctypes.data, buf.ctypes.data)
                if st:
                    raise RuntimeError(self.lib.pfft_emu_last_error().decode())
            return bi.value
        finally:
            self.lib.pfft_emu_plan_destroy(h)

    def nbytes(self, dims, kinds, real_type, permute=True, extra=()):
        N = len(dims)
        h = C.c_void_p()
        ne = len(extra)
        st = self.lib.pfft_emu_plan_create(
            C.byref(h), N, (C.c_int64 * N)(*dims), (C.c_int32 * N)(*kinds), real_type,
            int(permute), ne, (C.c_int64 * max(ne, 1))(*extra), 0)
        if st:
            raise RuntimeError(self.lib.pfft_emu_last_error().decode())
        try:
            bi, bo = C.c_size_t(), C.c_size_t()
            self.lib.pfft_emu_nbytes(h, C.byref(bi), C.byref(bo))
            return bi.value, bo.value
        finally:
            self.lib.pfft_emu_plan_destroy(h)

    def run_op(self, dims, kinds, real_type, direction, x, op, out=None, permute=True, extra=(),
               traversal=(0, 0)):
        """pfft_emu_execute_op: `op` is a ctypes pfft_spectral_op; `out` (uint8,
        optional) is the initial output for accumulate."""
        N = len(dims)
        h = C.c_void_p()
        ne = len(extra)
        st = self.lib.pfft_emu_plan_create(
            C.byref(h), N, (C.c_int64 * N)(*dims), (C.c_int32 * N)(*kinds), real_type,
            int(permute), ne, (C.c_int64 * max(ne, 1))(*extra), 0)
        if st:
            raise RuntimeError(self.lib.pfft_emu_last_error().decode())
        try:
            self.lib.pfft_emu_set_traversal(h, *traversal)
            bi, bo = C.c_size_t(), C.c_size_t()
            self.lib.pfft_emu_nbytes(h, C.byref(bi), C.byref(bo))
            nin, nout = (bi.value, bo.value) if direction in (0, 4) else (bo.value, bi.value)
            src = np.ascontiguousarray(x).reshape(-1)
            assert src.nbytes == nin, (src.nbytes, nin)
            res = np.zeros(nout, dtype=np.uint8) if out is None else out.copy()
            st = self.lib.pfft_emu_execute_op(h, direction, C.addressof(op), src.ctypes.data,
                                              res.ctypes.data)
            if st:
                raise RuntimeError(self.lib.pfft_emu_last_error().decode())
            return res
        finally:
            self.lib.pfft_emu_plan_destroy(h)
