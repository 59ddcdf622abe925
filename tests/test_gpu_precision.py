"""The precision extension: radix-16 stages of 1024-point c2c lines on MFMA,
split-FP16 (hi + lo) or BF16x3, FP32 accumulation.  Errors against the FP64
oracle must stay within 1e-6 relative L2; the plain FP32 path within
5e-7 log2(N)."""
import numpy as np
import pytest

from oracle import pipeline as OP
from oracle import plan as OM

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FLAGS = {"fp32": 0, "split16": 4, "bf16x3": 8}


def _case(pf, dims, flag, permute=True):
    T = pf.Transforms
    plan = pf.PencilFFTPlan(dims, T.FFT(), (1,), None, T.Float32, permute_dims=permute,
                            flags=FLAGS[flag])
    ks = plan.kernels(0) + plan.kernels(2)
    u = pf.allocate_input(plan)
    g = torch.Generator(device="cuda").manual_seed(4)
    u.parent.copy_(torch.complex(torch.randn(u.parent.shape, generator=g, device="cuda"),
                                 torch.randn(u.parent.shape, generator=g, device="cuda")))
    v = plan * u
    w = plan.ldiv(v)
    torch.cuda.synchronize()
    x = u.parent.cpu().numpy().astype(np.complex128).reshape(-1)
    meta = OM.PlanMeta(dims, ("FFT",), "Float64", permute)
    ref = OP.memory_image(OP.forward(meta, OP.from_image(x, meta.mem_dims("in")), workers=16))
    err_f = OP.rel_l2(v.parent.cpu().numpy().reshape(-1), ref)
    err_b = OP.rel_l2(w.parent.cpu().numpy().reshape(-1), x)
    return ks, err_f, err_b


@pytest.mark.parametrize("flag", ["split16", "bf16x3"])
@pytest.mark.parametrize("dims,permute", [((1024, 32, 16), True), ((16, 1024, 8), True),
                                          ((1024, 24, 16), False), ((1024, 1024), True)])
def test_mfma_variants(pf, gpu, dims, flag, permute):
    ks, err_f, err_b = _case(pf, dims, flag, permute)
    print(dims, flag, ks, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert any("mfma" in k for k in ks), ks
    assert err_f <= 1e-6 and err_b <= 1e-6


def test_plain_fp32_reference_point(pf, gpu):
    dims = (1024, 32, 16)
    ks, err_f, err_b = _case(pf, dims, "fp32")
    print("fp32", ks, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert not any("mfma" in k for k in ks)
    assert err_f <= 5e-7 * np.log2(np.prod(dims))
