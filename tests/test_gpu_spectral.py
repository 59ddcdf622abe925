"""Spectral operators fused into the passes on the GPU (pfft_execute_op
through the C ABI), on shapes whose carrying passes run the specialised
operator variants, against numpy multipliers on the FP64 oracle's spectra;
transposes by the dot-product identity; the batched gradient workload built
on them is covered in test_gpu_autodiff.py."""
import numpy as np
import pytest

from oracle import pipeline as OP
from tests.cases import CODE_NAMES, meta_for, random_logical, tol_for
from tests.test_spectral_ops import multiplier

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _fill(a, x):
    a.logical().copy_(torch.from_numpy(np.ascontiguousarray(x)).to(a.parent.dtype))


def _image(a):
    return a.parent.detach().cpu().numpy().reshape(-1)


SHAPES = [(64, 32, 16), (128, 64, 32), (256, 128, 64), (96, 64, 48), (512, 64, 32),
          (1024, 32, 16)]
OPS = [(1, 0), (1, 1), (1, 2), (2, 0), (3, 0), (4, 0)]


@pytest.mark.parametrize("real_type", [0, 1], ids=["F32", "F64"])
@pytest.mark.parametrize("dims", SHAPES, ids=lambda d: "x".join(map(str, d)))
def test_operators_forward_ldiv(pf, gpu, dims, real_type):
    # (1024, 32, 16): the r2c of 1024 reals, a 512-point complex FFT per line
    T = pf.Transforms
    S = pf.spectral
    meta, _ = meta_for(T, dims, (T.RFFT(),), CODE_NAMES[real_type])
    plan = pf.PencilFFTPlan(dims, T.RFFT(), (1,), None, real_type)
    u = pf.allocate_input(plan)
    v = pf.allocate_output(plan)
    w = pf.allocate_input(plan)
    rng = np.random.default_rng(4)
    x = random_logical(rng, dims, "Float64")
    _fill(u, x)
    xd = u.logical().cpu().numpy().astype(np.float64)
    y = OP.forward(meta, OP.to_memory(xd, meta.perm_in))
    tol = tol_for(CODE_NAMES[real_type], dims)
    for op, dim in OPS:
        mm = OP.to_memory(multiplier(op, dim, dims, True), meta.perm_out)
        S.mul_(v, plan, u, op, dim=dim, alpha=0.5)
        torch.cuda.synchronize()
        err_f = OP.rel_l2(_image(v), OP.memory_image(0.5 * mm * y))
        # ldiv with the operator from the exact spectrum
        v.parent.copy_(torch.from_numpy(np.ascontiguousarray(np.transpose(y)))
                       .to(v.parent.dtype))
        S.ldiv_(w, plan, v, op, dim=dim, alpha=2.0)
        torch.cuda.synchronize()
        want = 2.0 * OP.backward(meta, mm * y, scaled=True)
        err_b = OP.rel_l2(_image(w), OP.memory_image(np.real(want)))
        print(dims, real_type, op, dim, "fwd %.2e ldiv %.2e" % (err_f, err_b))
        assert err_f <= tol and err_b <= 10 * tol, (op, dim, err_f, err_b)


@pytest.mark.parametrize("dims", [(64, 32, 16), (256, 128, 64), (512, 64, 32)],
                         ids=lambda d: "x".join(map(str, d)))
def test_operator_transposes_and_accumulate(pf, gpu, dims):
    T = pf.Transforms
    S = pf.spectral
    L = pf._lib
    plan = pf.PencilFFTPlan(dims, T.RFFT(), (1,), None, T.Float64)
    g = torch.Generator(device="cuda").manual_seed(6)
    x = pf.allocate_input(plan)
    x.parent.normal_(generator=g)
    yb = pf.allocate_output(plan)
    yb.parent.copy_(torch.complex(torch.randn(yb.parent.shape, generator=g, device="cuda",
                                              dtype=torch.float64),
                                  torch.randn(yb.parent.shape, generator=g, device="cuda",
                                              dtype=torch.float64)))
    Ax = pf.allocate_output(plan)
    ATy = pf.allocate_input(plan)
    for op, dim in [(1, 0), (1, 2), (3, 0), (4, 0)]:
        S.mul_(Ax, plan, x, op, dim=dim, alpha=1.5)
        S.mul_adjoint_(ATy, plan, yb, op, dim=dim, alpha=1.5)
        lhs = float(torch.sum((torch.conj(yb.parent) * Ax.parent).real))
        rhs = float(torch.sum(ATy.parent * x.parent))
        assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1.0), (op, dim, lhs, rhs)
        By = pf.allocate_input(plan)
        BTx = pf.allocate_output(plan)
        S.bmul_(By, plan, yb, op, dim=dim)
        S.bmul_adjoint_(BTx, plan, x, op, dim=dim)
        lhs = float(torch.sum(x.parent * By.parent))
        rhs = float(torch.sum((torch.conj(BTx.parent) * yb.parent).real))
        assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1.0), (op, dim, lhs, rhs)
    # accumulate: sum_j i k_j F(u_j) = three accumulating forwards
    us = [pf.allocate_input(plan) for _ in range(3)]
    for a in us:
        a.parent.normal_(generator=g)
    acc = pf.allocate_output(plan)
    for j in range(3):
        S.mul_(acc, plan, us[j], S.DERIVATIVE, dim=j, accumulate=j > 0)
    ref = pf.allocate_output(plan)
    tmp = pf.allocate_output(plan)
    ref.parent.zero_()
    for j in range(3):
        S.mul_(tmp, plan, us[j], S.DERIVATIVE, dim=j)
        ref.parent.add_(tmp.parent)
    assert torch.allclose(acc.parent, ref.parent, rtol=1e-12, atol=1e-9)
    assert L.OP_DERIVATIVE == S.DERIVATIVE


def test_gradient_of_a_mode(pf, gpu):
    """d/dx sin(3x) cos(2y) = 3 cos(3x) cos(2y), exactly, from the fused
    derivative of the spectrum (the reference's gradient example,
    docs/examples/gradient.jl)."""
    T = pf.Transforms
    S = pf.spectral
    n = 64
    plan = pf.PencilFFTPlan((n, n, n), T.RFFT(), (1,), None, T.Float64)
    xs = torch.arange(n, dtype=torch.float64, device="cuda") * (2 * np.pi / n)
    X, Y, Z = torch.meshgrid(xs, xs, xs, indexing="ij")
    u = pf.allocate_input(plan)
    u.logical().copy_(torch.sin(3 * X) * torch.cos(2 * Y))
    uh = plan * u
    d = pf.allocate_input(plan)
    S.ldiv_(d, plan, uh, S.DERIVATIVE, dim=0)
    assert torch.allclose(d.logical(), 3 * torch.cos(3 * X) * torch.cos(2 * Y), atol=1e-12)
    S.ldiv_(d, plan, uh, S.DERIVATIVE, dim=1)
    assert torch.allclose(d.logical(), -2 * torch.sin(3 * X) * torch.sin(2 * Y), atol=1e-12)
    S.ldiv_(d, plan, uh, S.LAPLACIAN)
    assert torch.allclose(d.logical(), -13 * torch.sin(3 * X) * torch.cos(2 * Y), atol=1e-10)
    S.ldiv_(d, plan, uh, S.INV_LAPLACIAN)
    assert torch.allclose(d.logical(), -torch.sin(3 * X) * torch.cos(2 * Y) / 13, atol=1e-13)


@pytest.mark.parametrize("dims,real_type", [((256, 64, 32), 0), ((96, 64, 48), 0),
                                            ((512, 32, 16), 1), ((1024, 32, 16), 0)],
                         ids=["256F32", "96F32", "512F64", "1024F32"])
def test_every_operator_variant(pf, gpu, dims, real_type):
    """Each operator-carrying variant the tuner may pick (forced with
    PFFT_FLAG_CANDIDATE, tiles walked both ways) on both sides: mul! with the
    derivative and ldiv! with the inverse Laplacian, against the oracle."""
    T = pf.Transforms
    S = pf.spectral
    L = pf._lib
    meta, _ = meta_for(T, dims, (T.RFFT(),), CODE_NAMES[real_type])
    rng = np.random.default_rng(9)
    x = random_logical(rng, dims, "Float64")
    tol = tol_for(CODE_NAMES[real_type], dims)
    m1 = OP.to_memory(multiplier(1, 2, dims, True), meta.perm_out)
    m3 = OP.to_memory(multiplier(3, 0, dims, True), meta.perm_out)
    for k in range(24):
        # candidates 0..11 in both tile directions, the reversed ones with
        # streaming stores
        flags = L.FLAG_CANDIDATE(k % 12) | (
            L.FLAG_REVERSE_TILES | L.FLAG_STREAM_STORES if k >= 12 else 0)
        plan = pf.PencilFFTPlan(dims, T.RFFT(), (1,), None, real_type, flags=flags)
        u = pf.allocate_input(plan)
        v = pf.allocate_output(plan)
        w = pf.allocate_input(plan)
        _fill(u, x)
        xd = u.logical().cpu().numpy().astype(np.float64)
        y = OP.forward(meta, OP.to_memory(xd, meta.perm_in))
        S.mul_(v, plan, u, S.DERIVATIVE, dim=2)
        torch.cuda.synchronize()
        err_f = OP.rel_l2(_image(v), OP.memory_image(m1 * y))
        v.parent.copy_(torch.from_numpy(np.ascontiguousarray(np.transpose(y)))
                       .to(v.parent.dtype))
        S.ldiv_(w, plan, v, S.INV_LAPLACIAN)
        torch.cuda.synchronize()
        want = np.real(OP.backward(meta, m3 * y, scaled=True))
        err_b = OP.rel_l2(_image(w), OP.memory_image(want))
        print(k, dims, "fwd %.2e ldiv %.2e" % (err_f, err_b))
        assert err_f <= tol and err_b <= 10 * tol, (k, err_f, err_b)


def test_taylor_green_fused_curl_divergence(pf, gpu):
    """test/taylor_green.jl:95-150 with the fused operators: zero divergence
    and the analytic vorticity, on the reference's (4 pi, 2 pi, 2 pi) box."""
    T = pf.Transforms
    S = pf.spectral
    dims = (16, 8, 16)
    box = (4 * np.pi, 2 * np.pi, 2 * np.pi)
    scale = tuple(2 * np.pi / b for b in box)
    u0, k0 = 3.0, 2.0
    plan = pf.PencilFFTPlan(dims, T.RFFT(), (1, 1), None, T.Float64)
    xs = [np.arange(n) * (b / n) for n, b in zip(dims, box)]
    x, y, z = np.meshgrid(*xs, indexing="ij")
    fields = [u0 * np.sin(k0 * x) * np.cos(k0 * y) * np.cos(k0 * z),
              -u0 * np.cos(k0 * x) * np.sin(k0 * y) * np.cos(k0 * z),
              np.zeros(dims)]
    u = pf.allocate_input(plan, pf.Val(3))
    for a, f in zip(u, fields):
        _fill(a, f)
    div = S.divergence_(pf.allocate_output(plan), plan, u, scale)
    torch.cuda.synchronize()
    assert float(torch.sum(torch.abs(div.parent) ** 2)) / np.prod(dims) <= 1e-16
    wF = S.curl_(pf.allocate_output(plan, pf.Val(3)), plan, u, scale)
    w = [plan.ldiv(a) for a in wF]
    torch.cuda.synchronize()
    wt = [-u0 * k0 * np.cos(k0 * x) * np.sin(k0 * y) * np.sin(k0 * z),
          -u0 * k0 * np.sin(k0 * x) * np.cos(k0 * y) * np.sin(k0 * z),
          2 * u0 * k0 * np.sin(k0 * x) * np.sin(k0 * y) * np.cos(k0 * z)]
    err = sum(np.sum((a.logical().cpu().numpy() - b) ** 2) for a, b in zip(w, wt))
    assert err <= 1e-16, err
    # gradient of u_x from its spectrum: d_x u_x = u0 k0 cos cos cos
    g = S.gradient_([pf.allocate_input(plan) for _ in range(3)], plan, plan * u[0], scale)
    torch.cuda.synchronize()
    gx = u0 * k0 * np.cos(k0 * x) * np.cos(k0 * y) * np.cos(k0 * z)
    assert np.max(np.abs(g[0].logical().cpu().numpy() - gx)) <= 1e-12


@pytest.mark.parametrize("dims,real_type,extra", [((16, 12, 10), 1, ()), ((32, 24, 20), 0, ()),
                                                  ((12, 10, 8), 1, (2,))],
                         ids=["16x12x10-F64", "32x24x20-F32", "12x10x8-F64-extra2"])
def test_project_divergence_free(pf, gpu, dims, real_type, extra):
    """project_divergence_free! (docs/examples/navier_stokes.jl:351-366) on
    random spectra of real fields: the numpy formula on the same input, zero
    divergence after, idempotence, and a divergence-free field unchanged."""
    T = pf.Transforms
    S = pf.spectral
    plan = pf.PencilFFTPlan(dims, T.RFFT(), (1, 1), None, real_type, extra_dims=extra)
    box = (2 * np.pi, 4 * np.pi, 2 * np.pi)
    scale = tuple(2 * np.pi / b for b in box)
    g = torch.Generator(device="cuda").manual_seed(3)
    us = [pf.allocate_input(plan) for _ in range(3)]
    for a in us:
        a.parent.normal_(generator=g)
    uh = [plan * a for a in us]
    torch.cuda.synchronize()
    before = [a.logical().cpu().numpy().astype(np.complex128) for a in uh]
    S.project_divergence_free_(uh, plan, scale)
    torch.cuda.synchronize()
    got = [a.logical().cpu().numpy().astype(np.complex128) for a in uh]
    ks = [np.arange(dims[0] // 2 + 1) * scale[0],
          np.fft.fftfreq(dims[1], 1.0 / dims[1]) * scale[1],
          np.fft.fftfreq(dims[2], 1.0 / dims[2]) * scale[2]]
    K = [k.reshape(k.shape + (1,) * len(extra)) for k in np.meshgrid(*ks, indexing="ij")]
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    kdotu = sum(Ki * ui for Ki, ui in zip(K, before)) / np.where(k2 > 0, k2, 1.0)
    want = [np.where(k2 > 0, ui - Ki * kdotu, ui) for Ki, ui in zip(K, before)]
    tol = 1e-13 if real_type == 1 else 1e-6
    norm = np.sqrt(sum(np.sum(np.abs(w) ** 2) for w in want))
    assert np.sqrt(sum(np.sum(np.abs(a - b) ** 2) for a, b in zip(got, want))) <= tol * norm
    div = sum(Ki * ui for Ki, ui in zip(K, got))
    assert np.sqrt(np.sum(np.abs(div) ** 2)) <= tol * norm * max(np.max(np.abs(ks[1])), 1)
    # idempotent
    S.project_divergence_free_(uh, plan, scale)
    torch.cuda.synchronize()
    again = [a.logical().cpu().numpy().astype(np.complex128) for a in uh]
    assert np.sqrt(sum(np.sum(np.abs(a - b) ** 2) for a, b in zip(again, got))) <= tol * norm
    with pytest.raises(pf.ArgumentError):
        S.project_divergence_free_(uh[:2], plan, scale)
