"""The GPU workloads at their full sizes, checked through size-independent
properties and closed forms (the reference's own checks are a forward
transform against serial FFTW and a round trip, test/transforms.jl:258-299):

  * 512^3 ComplexF32 FFT + BFFT: against the FP64 oracle on N(0, 1) data,
    and against the exact spectrum of a separable input;
  * 1024^3 ComplexF32 FFT + BFFT with the plain FP32 radix stages and the
    split-FP16 / BF16x3 MFMA radix-16 stages: the separable input's exact
    spectrum, the round trip, and Parseval on N(0, 1) data;
  * the mixed workload (RFFT, FFT, R2R(REDFT10)) on its 384x384x256
    Fourier-Chebyshev field: the closed-form spectrum of the noiseless
    field, and the oracle on the field with noise.

Separable input u(x, y, z) = a(x) b(y) c(z) has the 3D DFT
U(kx, ky, kz) = A(kx) B(ky) C(kz), the outer product of three FP64 1D DFTs.
a, b, c take small Gaussian-integer values, so every product is exact in
Float32 and the reference spectrum is exact.  The comparison runs on the GPU
in float64, one slab at a time.

Tolerances: Float32 5e-7 log2(N) (N the total size); the MFMA variants
1e-6; Parseval's relative defect twice the tolerance of the transform.
"""
import math

import numpy as np
import pytest

from oracle import pipeline as OP
from oracle import plan as OM

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FLAGS = {"fp32": 0, "split16": 4, "bf16x3": 8}


def _gauss_int(rng, n, lim=8):
    return (rng.integers(-lim, lim + 1, n) + 1j * rng.integers(-lim, lim + 1, n)).astype(
        np.complex128)


def _separable(dims, seed):
    rng = np.random.default_rng(seed)
    return [_gauss_int(rng, n) for n in dims]


def _fill_separable(u, vecs):
    """u.parent is indexed [z, y, x] (memory dims (X, Y, Z) reversed)."""
    a, b, c = (torch.from_numpy(v).to("cuda") for v in vecs)
    p = u.parent
    for z0 in range(0, p.shape[0], 64):
        z1 = min(p.shape[0], z0 + 64)
        blk = c[z0:z1, None, None] * b[None, :, None] * a[None, None, :]
        p[z0:z1].copy_(blk.to(p.dtype))


def _rel_err_separable(t, vecs, order):
    """||t - outer|| / ||outer|| with t indexed by `order` (a permutation of
    the three vectors' axes), in float64 on the GPU, slab by slab."""
    v = [torch.from_numpy(x).to("cuda") for x in vecs]
    p0, p1, p2 = (v[i] for i in order)
    num = 0.0
    den = 0.0
    for i0 in range(0, t.shape[0], 32):
        i1 = min(t.shape[0], i0 + 32)
        ref = p0[i0:i1, None, None] * p1[None, :, None] * p2[None, None, :]
        d = t[i0:i1].to(torch.complex128) - ref
        num += float(torch.sum(d.real ** 2 + d.imag ** 2))
        den += float(torch.sum(ref.real ** 2 + ref.imag ** 2))
    return math.sqrt(num / den)


def _sqnorm(t):
    s = 0.0
    for i0 in range(0, t.shape[0], 64):
        x = t[i0:i0 + 64].to(torch.complex128)
        s += float(torch.sum(x.real ** 2 + x.imag ** 2))
    return s


def _c2c_full(pf, edge, flag):
    T = pf.Transforms
    dims = (edge,) * 3
    n = edge ** 3
    tol = 1e-6 if flag != "fp32" else 5e-7 * math.log2(n)
    plan = pf.PencilFFTPlan(dims, T.FFT(), (1, 1), None, T.Float32, flags=FLAGS[flag])
    ks = plan.kernels(0) + plan.kernels(2)
    print(edge, flag, ks, flush=True)
    if flag != "fp32":
        assert all("mfma" in k for k in ks if k.startswith("k_fast<f32,1024")), ks
    u = pf.allocate_input(plan)
    v = pf.allocate_output(plan)
    vecs = _separable(dims, seed=edge)
    _fill_separable(u, vecs)
    pf.mul_(v, plan, u)
    torch.cuda.synchronize()
    spec = [np.fft.fft(x) for x in vecs]
    # output memory dims (Z, Y, X): parent indexed [kx, ky, kz]
    err_f = _rel_err_separable(v.parent, spec, (0, 1, 2))
    pf.ldiv_(u, plan, v)
    torch.cuda.synchronize()
    # input parent indexed [z, y, x]
    err_b = _rel_err_separable(u.parent, vecs, (2, 1, 0))
    # Parseval on N(0, 1) data: sum |U|^2 = N sum |u|^2
    g = torch.Generator(device="cuda").manual_seed(5)
    u.parent.real.normal_(generator=g)
    u.parent.imag.normal_(generator=g)
    su = _sqnorm(u.parent)
    pf.mul_(v, plan, u)
    torch.cuda.synchronize()
    sv = _sqnorm(v.parent)
    parseval = abs(sv / (n * su) - 1.0)
    print("%d^3 C32 %s: separable fwd %.2e, round trip %.2e, Parseval defect %.2e"
          % (edge, flag, err_f, err_b, parseval), flush=True)
    assert err_f <= tol, err_f
    assert err_b <= tol, err_b
    assert parseval <= 2 * tol, parseval
    del plan, u, v
    torch.cuda.empty_cache()


@pytest.mark.parametrize("flag", ["fp32", "split16", "bf16x3"])
def test_c2c_1024_full_size(pf, gpu, flag):
    """The precision workload: 1024^3 ComplexF32 FFT + BFFT (8.6 GB arrays)."""
    _c2c_full(pf, 1024, flag)


def test_c2c_512_full_size_separable(pf, gpu):
    _c2c_full(pf, 512, "fp32")


def test_c2c_512_full_size_oracle(pf, gpu):
    """512^3 ComplexF32 FFT + BFFT on N(0, 1) data against the FP64 oracle
    (per-axis bound 5e-7 log2(512) = 4.5e-6, stricter than 5e-7 log2(N))."""
    T = pf.Transforms
    dims = (512, 512, 512)
    plan = pf.PencilFFTPlan(dims, T.FFT(), (1, 1), None, T.Float32)
    u = pf.allocate_input(plan)
    g = torch.Generator(device="cuda").manual_seed(42)
    u.parent.real.normal_(generator=g)
    u.parent.imag.normal_(generator=g)
    v = plan * u
    w = plan.ldiv(v)
    torch.cuda.synchronize()
    x = u.parent.cpu().numpy().astype(np.complex128).reshape(-1)
    err_b = OP.rel_l2(w.parent.cpu().numpy().reshape(-1), x)
    meta = OM.PlanMeta(dims, ("FFT",), "Float64")
    ref = OP.forward(meta, OP.from_image(x, dims), workers=16)
    err_f = OP.rel_l2(v.parent.cpu().numpy().reshape(-1), OP.memory_image(ref))
    print("512^3 C32 FFT vs oracle: fwd %.2e rt %.2e" % (err_f, err_b), flush=True)
    assert err_f <= 4.5e-6 and err_b <= 4.5e-6


def test_mixed_workload_closed_form(pf, gpu):
    """The noiseless Fourier-Chebyshev field's spectrum under (RFFT, FFT,
    R2R(REDFT10)) is a sum of delta peaks in closed form; the round trip
    through (BRFFT, BFFT, REDFT01) returns the field."""
    T = pf.Transforms
    W = pf.workloads
    dims = (384, 384, 256)
    plan = pf.PencilFFTPlan(dims, (T.RFFT(), T.FFT(), T.R2R(T.REDFT10)), (1, 1), None,
                            T.Float32)
    ks = plan.kernels(0) + plan.kernels(2)
    assert all(k.startswith("k_fast") for k in ks), ks
    u = pf.allocate_input(plan)
    field = W.fourier_chebyshev_field(dims, noise=0.0, seed=3)
    u.logical().copy_(field)
    v = plan * u
    torch.cuda.synchronize()
    S = W.fourier_chebyshev_spectrum(dims, seed=3)
    got = v.logical().cpu().numpy().astype(np.complex128)
    err_f = np.linalg.norm(got - S) / np.linalg.norm(S)
    w = plan.ldiv(v)
    torch.cuda.synchronize()
    err_b = float(torch.linalg.norm((w.logical() - field).double()) /
                  torch.linalg.norm(field.double()))
    tol = 5e-7 * math.log2(float(np.prod(dims)))
    print("mixed closed form: fwd %.2e rt %.2e" % (err_f, err_b), flush=True)
    assert err_f <= tol and err_b <= tol
