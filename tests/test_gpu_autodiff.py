"""Differentiation through the HIP pipeline (the batched + differentiable
workload): adjoint kernels, autograd, finite differences, batched fields in
one hipGraph.

References: torch's own FFT on the CPU in float64 (pocketfft, a library
independent of this build) with its autograd gives the reference gradient;
central differences with h = 1e-3 check it directionally (the loss is
quadratic, so the central difference is exact up to rounding)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _inner(a, b):
    return float(torch.sum((torch.conj(a) * b).real if a.is_complex() else a * b))


@pytest.mark.parametrize("dims", [(64, 32, 16), (15, 7, 5), (42, 24, 16), (256, 64, 32),
                                  (96, 64, 48)], ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("tname", ["RFFT", "BRFFT"])
@pytest.mark.parametrize("real_type", [0, 1], ids=["F32", "F64"])
def test_adjoint_dot_product(pf, gpu, dims, tname, real_type):
    T = pf.Transforms
    L = pf._lib
    tr = T.RFFT() if tname == "RFFT" else T.BRFFT(dims[-1] % 2 == 0)
    plan = pf.PencilFFTPlan(dims, tr, (1,), None, real_type)
    g = torch.Generator(device="cuda").manual_seed(3)

    def rnd(a):
        if a.parent.is_complex():
            a.parent.copy_(torch.complex(torch.randn(a.parent.shape, generator=g, device="cuda"),
                                         torch.randn(a.parent.shape, generator=g,
                                                     device="cuda")))
        else:
            a.parent.normal_(generator=g)
        return a

    x = rnd(pf.allocate_input(plan))
    yb = rnd(pf.allocate_output(plan))
    Ax = pf.allocate_output(plan)
    ATy = pf.allocate_input(plan)
    plan._execute(L.FORWARD, x, Ax)
    plan._execute(L.ADJOINT_FORWARD, yb, ATy)
    lhs, rhs = _inner(yb.parent.double() if not yb.parent.is_complex() else yb.parent.cdouble(),
                      Ax.parent.double() if not Ax.parent.is_complex() else Ax.parent.cdouble()), \
        _inner(ATy.parent.double() if not ATy.parent.is_complex() else ATy.parent.cdouble(),
               x.parent.double() if not x.parent.is_complex() else x.parent.cdouble())
    tol = 1e-12 if real_type == 1 else 2e-5
    assert abs(lhs - rhs) <= tol * max(abs(lhs), 1.0), (lhs, rhs)
    xb = rnd(pf.allocate_input(plan))
    By = pf.allocate_input(plan)
    BTx = pf.allocate_output(plan)
    plan._execute(L.BACKWARD, yb, By)
    plan._execute(L.ADJOINT_BACKWARD, xb, BTx)
    c = lambda t: t.cdouble() if t.is_complex() else t.double()
    lhs, rhs = _inner(c(xb.parent), c(By.parent)), _inner(c(BTx.parent), c(yb.parent))
    assert abs(lhs - rhs) <= tol * max(abs(lhs), 1.0), (lhs, rhs)


def ref_energy_grad(u_logical):
    """sum_j sum_x (d_j u)^2 and its gradient with torch.fft on the CPU in
    float64; u_logical[x, y, z, c]."""
    u = u_logical.detach().cpu().double().clone().requires_grad_()
    e = u.shape[0]
    uh = torch.fft.rfftn(u, dim=(1, 2, 0))                     # r2c along x
    X = e // 2 + 1
    kx = torch.arange(X, dtype=torch.float64).view(X, 1, 1, 1)
    kf = torch.fft.fftfreq(e, 1.0 / e, dtype=torch.float64)
    ks = [kx, kf.view(1, e, 1, 1), kf.view(1, 1, e, 1)]
    loss = 0
    for k in ks:
        g = torch.fft.irfftn(1j * k * uh, s=(e, e, e), dim=(1, 2, 0))
        loss = loss + torch.sum(g ** 2)
    loss.backward()
    return float(loss), u.grad


def test_autograd_matches_torch_reference(pf, gpu):
    from importlib import import_module
    AD = import_module(pf.__name__ + ".autodiff")
    T = pf.Transforms
    e = 64
    for real_type, tol in ((1, 1e-11), (0, 2e-5)):
        dt = torch.float64 if real_type == 1 else torch.float32
        plan_u = pf.PencilFFTPlan((e, e, e), T.RFFT(), (1,), None, real_type, extra_dims=(3,))
        plan_g = pf.PencilFFTPlan((e, e, e), T.RFFT(), (1,), None, real_type, extra_dims=(3,))
        gen = torch.Generator(device="cpu").manual_seed(1)
        u_log = torch.randn((e, e, e, 3), generator=gen, dtype=torch.float64)
        u = u_log.permute(3, 2, 1, 0).contiguous().to(dt).cuda().requires_grad_()
        uh = AD.forward(plan_u, u)                           # (c, x', y, z)
        X = e // 2 + 1
        kx = torch.arange(X, device="cuda", dtype=dt).view(X, 1, 1)
        kf = torch.fft.fftfreq(e, 1.0 / e, device="cuda", dtype=dt)
        loss = 0
        for k in (kx, kf.view(1, e, 1), kf.view(1, 1, e)):
            gspec = 1j * k * uh
            g = AD.ldiv(plan_g, gspec)
            loss = loss + torch.sum(g.double() ** 2)
        loss.backward()
        grad = u.grad.permute(3, 2, 1, 0)                    # logical (x, y, z, c)
        ref_loss, ref_grad = ref_energy_grad(u_log.to(dt))
        assert abs(float(loss) - ref_loss) <= tol * ref_loss
        err = float(torch.linalg.norm(grad.cpu().double() - ref_grad) /
                    torch.linalg.norm(ref_grad))
        print("autograd grad rel err", real_type, err)
        assert err <= tol


def _model(pf, edges, real_type):
    from importlib import import_module
    B = import_module(pf.__name__ + ".batched")
    m = B.GradientEnergy(edges, real_type)
    B.taylor_green_fields(m, seed=5)
    return B, m


def test_batched_graph_replays(pf, gpu):
    """One hipGraph runs loss terms and gradient of every field; each replay,
    with eager work between replays, equals the eager evaluation, and the
    gradient matches the reference."""
    B, m = _model(pf, [64, 96, 64, 128], 0)
    m.run()
    torch.cuda.synchronize()
    loss0 = float(m.loss())
    grads0 = [m.grad(i).clone() for i in range(len(m.edges))]
    m.capture()
    for rep in range(3):
        for i in range(len(m.edges)):
            m.grad(i).zero_()           # eager work between replays
        m.replay()
        torch.cuda.synchronize()
        assert float(m.loss()) == pytest.approx(loss0, rel=1e-6)
        for i in range(len(m.edges)):
            assert torch.equal(m.grad(i), grads0[i]), (rep, i)
    for i in (0, 1, 3):
        ref_loss, ref_grad = ref_energy_grad(m.field(i))
        err = float(torch.linalg.norm(m.grad(i).cpu().double() - ref_grad) /
                    torch.linalg.norm(ref_grad))
        print("field", i, m.edges[i], "grad rel err", err)
        assert err <= 2e-5


def test_finite_differences(pf, gpu):
    """Central differences, h = 1e-3, along random directions (Float64)."""
    B, m = _model(pf, [64, 96], 1)
    m.run()
    torch.cuda.synchronize()
    gen = torch.Generator(device="cuda").manual_seed(9)
    base = [m.field(i).clone() for i in range(2)]
    grads = [m.grad(i).clone() for i in range(2)]
    h = 1e-3
    for trial in range(2):
        dirs = [torch.randn(b.shape, generator=gen, device="cuda", dtype=b.dtype) for b in base]
        vals = []
        for sgn in (1, -1):
            for i in range(2):
                m.field(i).copy_(base[i] + sgn * h * dirs[i])
            m.run()
            torch.cuda.synchronize()
            vals.append(float(m.loss()))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = sum(float(torch.sum(g * d)) for g, d in zip(grads, dirs))
        print("FD", fd, "adjoint", an)
        assert abs(fd - an) <= 1e-7 * abs(an)
    for i in range(2):
        m.field(i).copy_(base[i])


def test_batched_workload_full(pf, gpu):
    """The batched workload: 12 three-component Float32 fields, cube edges
    drawn by seed from {64, 96, 128, 192, 256, 320}, Taylor-Green data plus
    N(0, 0.01) noise, loss and gradient in one graph replay."""
    from importlib import import_module
    B = import_module(pf.__name__ + ".batched")
    edges = B.draw_edges(12, seed=42)
    m = B.GradientEnergy(edges, 0)
    B.taylor_green_fields(m, seed=0)
    m.capture()
    m.replay()
    torch.cuda.synchronize()
    ks = sorted(range(12), key=lambda i: edges[i])
    for i in (ks[0], ks[-1]):
        ref_loss, ref_grad = ref_energy_grad(m.field(i))
        err = float(torch.linalg.norm(m.grad(i).cpu().double() - ref_grad) /
                    torch.linalg.norm(ref_grad))
        print("edges", edges, "field", i, edges[i], "grad rel err %.2e" % err)
        assert err <= 5e-5
    assert math.isfinite(float(m.loss()))
