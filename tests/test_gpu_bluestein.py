"""Lines with a large prime factor on the device: the generic pass's
Bluestein (chirp-z) form against the FP64 oracle (relative L2, Float64 1e-13,
Float32 5e-7 log2 N), forward and back.  The reference accepts any length
(FFTW plans them); the odd sizes of test/rfft.jl are the model."""
import pytest

from tests.cases import CODE_NAMES, tol_for
from tests.test_gpu_parity import run_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims,trs_name,bs", [
    ((1031, 24, 6), "FFT", "bluestein4096"),     # prime line length
    ((2062, 16, 4), "RFFT", "bluestein4096"),    # 1031-point half-length FFT
    ((2 * 1031, 8), "FFT", "bluestein8192"),     # 2 x 1031 (Float64: direct, LDS)
    ((40, 134, 12), "FFT", "bluestein512"),      # 2 x 67 along dim 2
    ((67, 32, 8), "R2R:REDFT10", "bluestein256"),
    ((67, 32, 8), "R2R:RODFT01", "bluestein256"),
    ((67, 32, 8), "R2R:REDFT11", "bluestein512"),
])
def test_bluestein_parity(pf, gpu, dims, trs_name, bs, real_type):
    T = pf.Transforms
    if trs_name.startswith("R2R:"):
        trs = (T.R2R(getattr(T, trs_name[4:])),)
    else:
        trs = (getattr(T, trs_name)(),)
    err_f, err_b, meta = run_case(pf, dims, trs, real_type)
    tol = tol_for(CODE_NAMES[real_type], dims)
    print(trs_name, dims, CODE_NAMES[real_type], meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    if not (dims == (2 * 1031, 8) and real_type == 1):
        assert any(k.endswith("/" + bs) for k in meta.kernels), meta.kernels
    assert err_f <= tol, err_f
    assert err_b <= tol, err_b


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims,trs_name,npass", [
    ((16384, 64), "FFT", 3),         # 128 x 128 four-step split
    ((24, 16384, 4), "RFFT", 4),     # split on the second dim, after the r2c
    ((12288, 16), "BFFT", 3),        # 96 x 128
    ((64 * 1031, 4), "FFT", 3),      # 64 x 1031: the second half in the Bluestein form
])
def test_four_step_parity(pf, gpu, dims, trs_name, npass, real_type):
    """c2c lines too long for one LDS-resident pass: the four-step split
    (two passes, twiddle w_M^(j2 k1) between them) against the oracle."""
    T = pf.Transforms
    trs = (getattr(T, trs_name)(),)
    err_f, err_b, meta = run_case(pf, dims, trs, real_type)
    tol = tol_for(CODE_NAMES[real_type], dims)
    print(trs_name, dims, CODE_NAMES[real_type], meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert len(meta.kernels) == 2 * npass, meta.kernels
    if dims[0] == 64 * 1031:
        assert any(k.endswith("/bluestein4096") for k in meta.kernels), meta.kernels
    assert err_f <= tol, err_f
    assert err_b <= tol, err_b
