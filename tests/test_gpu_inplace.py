"""In-place plans on the GPU through the C ABI: ManyPencilArray /
ManyPencilArrayRFFT! buffers, results against the out-of-place plan and the
FP64 oracle, the reference's error contract for in-place plans
(test/transforms.jl:170-256) and its RFFT! tests (test/rfft.jl:194-251)."""
import numpy as np
import pytest

from oracle import pipeline as OP
from tests.cases import CODE_NAMES, expand, meta_for, random_logical, tol_for

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _fill(a, x):
    a.logical().copy_(torch.from_numpy(np.ascontiguousarray(x)).to(a.parent.dtype))


def _oop(pf, trs):
    L = pf._lib
    return tuple(pf.Transforms._from_code(t.code & ~L.INPLACE) for t in trs)


def run_inplace(pf, dims, trs, real_type, permute=True, extra=(), seed=0):
    T = pf.Transforms
    full = expand(T, trs, len(dims))
    oop = _oop(pf, full)
    plan = pf.PencilFFTPlan(dims, full, (1,), None, real_type, permute_dims=permute,
                            extra_dims=extra)
    plan_oop = pf.PencilFFTPlan(dims, oop, (1,), None, real_type, permute_dims=permute,
                                extra_dims=extra)
    assert plan.is_inplace() and not plan_oop.is_inplace()
    meta, _ = meta_for(T, dims, oop, CODE_NAMES[real_type], permute, extra)
    A = pf.allocate_input(plan)
    assert isinstance(A, pf.ManyPencilArray)
    rng = np.random.default_rng(seed)
    x = random_logical(rng, tuple(dims) + tuple(extra), meta.eltype_in)
    _fill(A.first(), x)
    u = pf.allocate_input(plan_oop)
    _fill(u, x)
    x_dev = A.first().logical().cpu().numpy().copy()
    assert plan * A is A                                   # in place
    v = plan_oop * u
    torch.cuda.synchronize()
    got = A.last().parent.cpu().numpy().reshape(-1)
    ref = OP.memory_image(OP.forward(meta, OP.to_memory(
        x_dev.astype(np.complex128 if np.iscomplexobj(x_dev) else np.float64), meta.perm_in)))
    err_f = OP.rel_l2(got, ref)
    same = OP.rel_l2(got, v.parent.cpu().numpy().reshape(-1))
    assert plan.ldiv(A) is A
    torch.cuda.synchronize()
    back = A.first().logical().cpu().numpy()
    err_b = OP.rel_l2(back.reshape(-1), x_dev.reshape(-1))
    return err_f, err_b, same, plan.kernels(0) + plan.kernels(2)


CASES = [
    ((16, 12, 6), "FFT!", 1), ((16, 12, 6), "FFT!", 0), ((15, 7, 5), "BFFT!", 1),
    ((16, 12, 6), "RFFT!", 1), ((15, 7, 5), "RFFT!", 0), ((16, 12, 6), "R2R!", 1),
    ((16, 12, 6), "mixed!", 0), ((11, 15), "FFT!", 1), ((9, 7, 5, 12), "RFFT!", 1),
    # the specialised kernels, in place and transposing between U and the scratch
    ((256, 128, 64), "FFT!", 0), ((512, 64, 32), "RFFT!", 0), ((256, 64, 32), "RFFT!", 1),
    ((1024, 32, 16), "FFT!", 0), ((384, 96, 64), "RFFT!", 0),
    # four-step split (16384 = 128 x 128) and Bluestein (1031) lines, in place
    ((16384, 32), "FFT!", 0), ((16384, 16), "FFT!", 1), ((1031, 24, 8), "FFT!", 1),
]


def _trs(T, name):
    return {"FFT!": (T.FFT_(),), "BFFT!": (T.BFFT_(),), "RFFT!": (T.RFFT_(),),
            "R2R!": (T.R2R_(T.REDFT10), T.R2R_(T.RODFT01), T.R2R_(T.DHT)),
            "mixed!": (T.FFT_(), T.NoTransform_(), T.BFFT_())}[name]


@pytest.mark.parametrize("dims,name,real_type", CASES,
                         ids=lambda c: "x".join(map(str, c)) if isinstance(c, tuple) else str(c))
def test_inplace_vs_oracle(pf, gpu, dims, name, real_type):
    err_f, err_b, same, ks = run_inplace(pf, dims, _trs(pf.Transforms, name), real_type, seed=7)
    tol = tol_for(CODE_NAMES[real_type], dims)
    print(dims, name, real_type, ks, "fwd %.2e rt %.2e vs oop %.2e" % (err_f, err_b, same))
    assert err_f <= tol and err_b <= tol and same <= tol


def test_inplace_extra_dims_and_noperm(pf, gpu):
    T = pf.Transforms
    for dims, permute, extra in [((16, 12, 6), True, (3,)), ((12, 10), False, (2,)),
                                 ((64, 32, 16), False, ())]:
        for trs in [(T.FFT_(),), (T.RFFT_(),)]:
            err_f, err_b, same, _ = run_inplace(pf, dims, trs, 1, permute, extra, seed=2)
            assert err_f <= 1e-13 and err_b <= 1e-13 and same <= 1e-13, (dims, trs)


def test_inplace_error_contract(pf, gpu):
    """test/transforms.jl:170-205 (test_transform, in-place)."""
    T = pf.Transforms
    plan = pf.PencilFFTPlan((16, 12, 6), T.FFT_(), (1,), None, T.Float64)
    plan_oop = pf.PencilFFTPlan((16, 12, 6), T.FFT(), (1,), None, T.Float64)
    vi = pf.allocate_input(plan)
    vo = pf.allocate_output(plan)
    assert type(vi) is type(vo)
    u, v = vi.first(), vi.last()
    with pytest.raises(pf.ArgumentError, match="incompatible with out-of-place plans"):
        plan_oop * vi
    with pytest.raises(pf.ArgumentError, match="aliased"):
        pf.mul_(v, plan_oop, u)
    vi_other = pf.allocate_input(plan)
    with pytest.raises(pf.ArgumentError, match="must be the same"):
        pf.mul_(vi_other, plan, vi)
    with pytest.raises(pf.ArgumentError, match="incompatible with in-place plans"):
        plan * u
    with pytest.raises(pf.ArgumentError, match="incompatible with in-place plans"):
        plan.ldiv(v)


def test_inplace_components(pf, gpu):
    """test/transforms.jl:237-253: collections of ManyPencilArrays."""
    T = pf.Transforms
    plan = pf.PencilFFTPlan((16, 12, 6), T.RFFT_(), (1,), None, T.Float64)
    for comp in ((pf.Val(3),), (3, 2)):
        vi = pf.allocate_input(plan, *comp)
        vi = list(vi) if isinstance(vi, tuple) else list(vi.reshape(-1, order="F"))
        g = torch.Generator(device="cuda").manual_seed(1)
        init = []
        for a in vi:
            a.first().logical().copy_(torch.randn(a.first().logical().shape, generator=g,
                                                  device="cuda", dtype=torch.float64))
            init.append(a.first().logical().clone())
        for a in vi:
            plan * a
        for a, x0 in zip(vi, init):
            assert not torch.allclose(a.first().logical(), x0)
        for a in vi:
            plan.ldiv(a)
        for a, x0 in zip(vi, init):
            assert torch.allclose(a.first().logical(), x0, atol=1e-12)


@pytest.mark.parametrize("dims", [(16, 12, 6), (41, 23, 15), (64, 32, 16)],
                         ids=lambda d: "x".join(map(str, d)))
def test_rfft_inplace_vs_rfft(pf, gpu, dims):
    """test/rfft.jl:194-251: RFFT! against RFFT on an impulse pair, ldiv!,
    bmul! on a random spectrum, then mul! again."""
    T = pf.Transforms
    ip = pf.PencilFFTPlan(dims, T.RFFT_(), (1,), None, T.Float64)
    op = pf.PencilFFTPlan(dims, T.RFFT(), (1,), None, T.Float64)
    A = pf.allocate_input(ip)
    x, xh = A.first(), A.last()
    v = pf.allocate_input(op)
    vh = pf.allocate_output(op)
    x.logical().zero_()
    v.logical().zero_()
    x.logical()[0, 0, 0] = 1.0
    x.logical()[1, 0, 0] = 2.0
    v.logical()[0, 0, 0] = 1.0
    v.logical()[1, 0, 0] = 2.0
    pf.mul_(A, ip, A)
    pf.mul_(vh, op, v)
    assert torch.allclose(xh.parent, vh.parent, atol=1e-8)
    pf.ldiv_(A, ip, A)
    pf.ldiv_(v, op, vh)
    assert torch.allclose(x.logical(), v.logical(), atol=1e-8)
    assert torch.allclose(x.logical()[:3, 0, 0],
                          torch.tensor([1.0, 2.0, 0.0], dtype=torch.float64, device="cuda"),
                          atol=1e-8)
    g = torch.Generator(device="cuda").manual_seed(42)
    # a random spectrum that is the transform of a real field
    w = pf.allocate_input(op)
    w.parent.normal_(generator=g)
    pf.mul_(vh, op, w)
    xh.parent.copy_(vh.parent)
    pf.bmul_(A, ip, A)
    pf.bmul_(v, op, vh)
    assert torch.allclose(x.logical(), v.logical(), atol=1e-8 * float(v.parent.abs().max()))
    pf.mul_(A, ip, A)
    pf.mul_(vh, op, v)
    assert torch.allclose(xh.parent, vh.parent, atol=1e-8 * float(vh.parent.abs().max()))
