"""Golden fixtures (tests/golden/, written by make_golden.py from the FP64
oracle at the reference's test shapes): the oracle still reproduces them, and
the kernels -- emulated on the CPU, and on the GPU -- match them within
  Float64: 1e-13 relative L2;  Float32: 5e-7 * log2(N)."""
import json
import os

import numpy as np
import pytest

from oracle import pipeline as OP
from oracle import plan as OM

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = ["g1_transform_list.npz", "g2_brfft_single_mode.npz", "g3_rfft_impulse.npz",
         "g4_taylor_green.npz", "g5_dimensionality.npz"]


def load_cases():
    meta = json.load(open(os.path.join(HERE, "g6_metadata.json")))
    cases = []
    for f in FILES:
        z = np.load(os.path.join(HERE, f))
        names = sorted({k.rsplit("/", 1)[0] for k in z.files})
        for n in names:
            cases.append((n, meta[n], z[n + "/in"], z[n + "/out"]))
    return cases


CASES = load_cases()
IDS = [c[0] for c in CASES]


def product_transforms(T, names):
    out = []
    for s in names:
        if s.startswith("R2R_"):
            out.append(T.R2R(getattr(T, s[4:])))
        elif s.startswith("BRFFT"):
            out.append(T.BRFFT(s == "BRFFT_even"))
        else:
            out.append(getattr(T, s)())
    return tuple(out)


def test_metadata_fixture_matches_oracle():
    meta = json.load(open(os.path.join(HERE, "g6_metadata.json")))
    assert len(meta) >= 30
    for name, m in meta.items():
        o = OM.PlanMeta(tuple(m["size_in"]), tuple(m["transforms"]), "Float64",
                        m["permute"], tuple(m["extra"]))
        assert list(o.size_out) == m["size_out"] and o.scale == m["scale"]
        assert [list(s["perm"]) for s in o.stages] == [s["perm"] for s in m["stages"]]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_reproduces(case):
    name, m, xin, yout = case
    o = OM.PlanMeta(tuple(m["size_in"]), tuple(m["transforms"]), "Float64", m["permute"])
    x = OP.from_image(xin, tuple(m["mem_in"]))
    for method in ("direct", "fast"):
        y = OP.memory_image(OP.forward(o, x, method=method))
        assert OP.rel_l2(y, yout) <= 1e-14, (name, method)


def _tol(real_type, dims):
    return 1e-13 if real_type == 1 else 5e-7 * max(np.log2(float(np.prod(dims))), 1.0)


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_emulated_matches_golden(pf, emu, case, real_type):
    name, m, xin, yout = case
    T = pf.Transforms
    trs = product_transforms(T, m["transforms"])
    cplx = m["eltype_in"].startswith("Complex")
    dt_in = (np.complex64 if cplx else np.float32) if real_type == 0 else \
        (np.complex128 if cplx else np.float64)
    ocplx = m["eltype_out"].startswith("Complex")
    dt_out = (np.complex64 if ocplx else np.float32) if real_type == 0 else \
        (np.complex128 if ocplx else np.float64)
    x = xin.astype(dt_in)
    out = emu.run(m["size_in"], [t.code for t in trs], real_type, 0, x, m["permute"]).view(dt_out)
    assert OP.rel_l2(out, yout) <= _tol(real_type, m["size_in"]), name


@pytest.mark.gpu
@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_gpu_matches_golden(pf, gpu, case, real_type):
    import torch
    name, m, xin, yout = case
    T = pf.Transforms
    trs = product_transforms(T, m["transforms"])
    plan = pf.PencilFFTPlan(tuple(m["size_in"]), trs, (1,), None, real_type,
                            permute_dims=m["permute"])
    u = pf.allocate_input(plan)
    src = torch.from_numpy(np.ascontiguousarray(xin)).to(u.parent.dtype)
    u.parent.view(-1).copy_(src.to(u.parent.device))
    v = plan * u
    torch.cuda.synchronize()
    got = v.parent.cpu().numpy().reshape(-1)
    assert OP.rel_l2(got, yout) <= _tol(real_type, m["size_in"]), name
