"""pfft_execute_noalloc (the entry point that allocates nothing, its
intermediates in the library's static device scratch) against pfft_execute
with a workspace, through the C ABI: bit-identical outputs on two calls in a
row, the second on a smaller plan, in every direction; and its error
contract (unprepared plan, plan larger than the scratch)."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _plan(L, dims, kinds, real_type, flags):
    h = C.c_void_p()
    n = len(dims)
    L.check(L.lib.pfft_plan_create(C.byref(h), n, (C.c_int64 * n)(*dims),
                                   (C.c_int32 * n)(*kinds), real_type, 1, 0,
                                   (C.c_int64 * 1)(), flags))
    return h


def _nbytes(L, h):
    bi, bo, bw = C.c_size_t(), C.c_size_t(), C.c_size_t()
    L.check(L.lib.pfft_plan_nbytes(h, C.byref(bi), C.byref(bo), C.byref(bw)))
    return bi.value, bo.value, bw.value


def _buf(nbytes, seed=None):
    t = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    if seed is None:
        t.fill_(float("nan"))
    else:
        t.uniform_(-1, 1, generator=torch.Generator(device="cuda").manual_seed(seed))
    return t


def _check_plan(L, dims, kinds, real_type, seed):
    h = _plan(L, dims, kinds, real_type, L.FLAG_NO_AUTOTUNE)
    try:
        L.check(L.lib.pfft_plan_prepare(h, None))
        nin, nout, nwork = _nbytes(L, h)
        assert nwork <= L.NOALLOC_SCRATCH_BYTES
        stream = torch.cuda.current_stream().cuda_stream
        work = _buf(max(nwork, 4))
        for direction, src_bytes, dst_bytes in ((L.FORWARD, nin, nout),
                                                (L.BACKWARD_SCALED, nout, nin),
                                                (L.BACKWARD, nout, nin),
                                                (L.ADJOINT_FORWARD, nout, nin),
                                                (L.ADJOINT_BACKWARD, nin, nout)):
            src = _buf(src_bytes, seed)
            ref = _buf(dst_bytes)
            got = _buf(dst_bytes)
            L.check(L.lib.pfft_execute(h, direction, C.c_void_p(src.data_ptr()),
                                       C.c_void_p(ref.data_ptr()), C.c_void_p(work.data_ptr()),
                                       nwork, C.c_void_p(stream)))
            L.check(L.lib.pfft_execute_noalloc(h, direction, C.c_void_p(src.data_ptr()),
                                               C.c_void_p(got.data_ptr()), C.c_void_p(stream)))
            torch.cuda.synchronize()
            assert torch.equal(got.view(torch.int32), ref.view(torch.int32)), (dims, direction)
        return nwork
    finally:
        L.lib.pfft_plan_destroy(h)


def test_noalloc_matches_execute_two_calls(pf, gpu):
    L = pf._lib
    # first call, then a second on a smaller plan; Float32 then Float64
    big = _check_plan(L, (128, 96, 64), (L.RFFT, L.FFT, L.FFT), L.F32, seed=1)
    small = _check_plan(L, (48, 32, 16), (L.RFFT, L.FFT, L.FFT), L.F32, seed=2)
    assert 0 < small < big
    _check_plan(L, (64, 40, 24), (L.FFT, L.FFT, L.FFT), L.F64, seed=3)
    _check_plan(L, (32, 24, 20), (L.RFFT, L.FFT, L.FFT), L.F64, seed=4)


def test_noalloc_error_contract(pf, gpu):
    L = pf._lib
    # larger than the scratch: 512^3 Float32 RFFT needs 2 x 539 MB
    h = _plan(L, (512, 512, 512), (L.RFFT, L.FFT, L.FFT), L.F32, L.FLAG_NO_AUTOTUNE)
    try:
        assert _nbytes(L, h)[2] > L.NOALLOC_SCRATCH_BYTES
        x = _buf(16)
        r = L.lib.pfft_execute_noalloc(h, L.FORWARD, C.c_void_p(x.data_ptr()),
                                       C.c_void_p(x.data_ptr() + 64), None)
        assert r == L.ERR_ARGUMENT
        assert b"static scratch" in L.lib.pfft_last_error()
    finally:
        L.lib.pfft_plan_destroy(h)
    # not prepared
    h = _plan(L, (16, 12, 8), (L.RFFT, L.FFT, L.FFT), L.F32, L.FLAG_NO_AUTOTUNE)
    try:
        nin, nout, _ = _nbytes(L, h)
        a, b = _buf(nin), _buf(nout)
        r = L.lib.pfft_execute_noalloc(h, L.FORWARD, C.c_void_p(a.data_ptr()),
                                       C.c_void_p(b.data_ptr()), None)
        assert r == L.ERR_ARGUMENT
        assert b"not prepared" in L.lib.pfft_last_error()
    finally:
        L.lib.pfft_plan_destroy(h)
