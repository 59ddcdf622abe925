"""The adjoint directions (PFFT_ADJOINT_FORWARD / PFFT_ADJOINT_BACKWARD) are
the transposes of mul! and bmul! for the real inner product
<a, b> = Re sum conj(a) b: the dot-product test <y_bar, A x> = <A^T y_bar, x>
fixes them independently of how they are computed (Hermitian weights w_k = 1
at DC and, for even lengths, Nyquist; 2 elsewhere)."""
import numpy as np
import pytest

from oracle import pipeline as OP
from tests.cases import NP_DTYPES, meta_for


def inner(a, b):
    return float(np.real(np.vdot(a.astype(np.complex128), b.astype(np.complex128))))


@pytest.mark.parametrize("dims", [(8,), (9,), (12, 10), (16, 12, 6), (15, 7, 5), (42, 24, 16)])
@pytest.mark.parametrize("tname", ["RFFT", "BRFFT", "FFT"])
def test_adjoint_dot_product_emulated(pf, emu, dims, tname):
    T = pf.Transforms
    if tname == "BRFFT":
        tr = T.BRFFT(2 * dims[-1] - 2 if dims[-1] % 2 == 0 else 2 * dims[-1] - 1)
        tr = T.BRFFT(dims[-1] % 2 == 0)
    else:
        tr = getattr(T, tname)()
    meta, full = meta_for(T, dims, (tr,), "Float64")
    kinds = [t.code for t in full]
    rng = np.random.default_rng(11)
    dt_in, dt_out = NP_DTYPES[meta.eltype_in], NP_DTYPES[meta.eltype_out]

    def rnd(shape, dt):
        x = rng.standard_normal(shape)
        if np.dtype(dt).kind == "c":
            x = x + 1j * rng.standard_normal(shape)
        return OP.memory_image(x).astype(dt)

    x = rnd(meta.mem_dims("in"), dt_in)
    yb = rnd(meta.mem_dims("out"), dt_out)
    Ax = emu.run(dims, kinds, 1, 0, x).view(dt_out)
    ATy = emu.run(dims, kinds, 1, 3, yb).view(dt_in)
    lhs, rhs = inner(yb, Ax), inner(ATy, x)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)
    # adjoint of bmul!: <x_bar, B y> = <B^T x_bar, y>
    xb = rnd(meta.mem_dims("in"), dt_in)
    By = emu.run(dims, kinds, 1, 1, yb).view(dt_in)
    BTx = emu.run(dims, kinds, 1, 4, xb).view(dt_out)
    lhs, rhs = inner(xb, By), inner(BTx, yb)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)


def test_adjoint_rejected_for_r2r(pf, emu):
    T = pf.Transforms
    with pytest.raises(RuntimeError, match="R2R"):
        emu.run((8, 6), [T.R2R(T.REDFT10).code] * 2, 1, 3, np.zeros(48))
