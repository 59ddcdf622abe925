"""Parity of the specialised power-of-two passes (k_fast) with the FP64
oracle, in every load/store form they take: contiguous lines, transposed
tiles on either side, r2c with a transposed store, c2r with a transposed
load, Float32 and Float64, permute_dims on and off, extra dims."""
import numpy as np
import pytest

from tests.test_gpu_parity import run_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHAPES = [(64, 64, 64), (128, 64, 256), (256, 128, 64), (512, 64, 128), (1024, 64, 32),
          (64, 1024, 16), (256, 512)]


@pytest.mark.parametrize("permute", [True, False], ids=["perm", "noperm"])
@pytest.mark.parametrize("real_type", [0, 1], ids=["F32", "F64"])
@pytest.mark.parametrize("tname", ["RFFT", "FFT", "BRFFT"])
@pytest.mark.parametrize("dims", SHAPES, ids=lambda d: "x".join(map(str, d)))
def test_fast_paths(pf, gpu, dims, tname, real_type, permute):
    T = pf.Transforms
    if tname == "BRFFT":
        # c2r along the last dim: complex input dims with last = n/2 + 1
        dims = tuple(dims[:-1]) + (dims[-1] // 2 + 1,)
        tr = T.BRFFT(2 * (dims[-1] - 1))
    else:
        tr = getattr(T, tname)()
    if real_type == 1 and max(dims) > 512:
        pytest.skip("Float64 fast kernels stop at 512")
    err_f, err_b, meta = run_case(pf, dims, (tr,), real_type, permute=permute, seed=3)
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    print(dims, tname, real_type, permute, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol


def test_fast_kernels_selected(pf, gpu):
    """The 512^3 Float32 RFFT pair runs entirely on the specialised kernels."""
    T = pf.Transforms
    plan = pf.PencilFFTPlan((512, 512, 512), T.RFFT(), (1, 1), None, T.Float32)
    fwd = plan.kernels(0)
    bwd = plan.kernels(2)
    print(fwd, bwd)
    assert all(k.startswith("k_fast") for k in fwd + bwd), (fwd, bwd)
