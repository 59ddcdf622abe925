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
    err_f, err_b, meta = run_case(pf, dims, (tr,), real_type, permute=permute, seed=3)
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    print(dims, tname, real_type, permute, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol
    c2c_dims = {"FFT": dims, "RFFT": dims[1:], "BRFFT": dims[:-1]}[tname]
    if real_type == 1 and 1024 in c2c_dims and permute:
        # Float64 c2c lines of 1024 points have specialised kernels (4-line
        # tiles on a transposed side); unpermuted, a pass strided on both
        # sides runs the generic kernel
        assert any(k.startswith("k_fast<f64,1024,c2c") for k in meta.kernels), meta.kernels


def test_fast_kernels_selected(pf, gpu):
    """The 512^3 Float32 RFFT pair runs entirely on the specialised kernels."""
    T = pf.Transforms
    plan = pf.PencilFFTPlan((512, 512, 512), T.RFFT(), (1, 1), None, T.Float32)
    fwd = plan.kernels(0)
    bwd = plan.kernels(2)
    print(fwd, bwd)
    assert all(k.startswith("k_fast") for k in fwd + bwd), (fwd, bwd)


MIXED_SHAPES = [(384, 96, 64), (192, 384, 32), (320, 160, 64), (96, 192, 48), (64, 320, 96),
                (768, 64, 32), (1536, 32, 16), (640, 64, 32), (1280, 32, 16), (32, 768, 48),
                (48, 640, 32)]


@pytest.mark.parametrize("permute", [True, False], ids=["perm", "noperm"])
@pytest.mark.parametrize("tname", ["RFFT", "FFT", "MIXED"])
@pytest.mark.parametrize("dims", MIXED_SHAPES, ids=lambda d: "x".join(map(str, d)))
def test_fast_mixed_radix(pf, gpu, dims, tname, permute):
    """Lengths 3*2^k and 5*2^k (the mixed and batched workloads' sizes, and the
    640- and 768-point lines, also as the half-length FFTs of 1280 and 1536
    reals) and the complex-data REDFT10/REDFT01 passes, Float32."""
    T = pf.Transforms
    if tname == "MIXED":
        trs = (T.RFFT(), T.FFT(), T.R2R(T.REDFT10))
    else:
        trs = (getattr(T, tname)(),)
    err_f, err_b, meta = run_case(pf, dims, trs, 0, permute=permute, seed=5)
    tol = 5e-7 * np.log2(float(np.prod(dims)))
    print(dims, tname, permute, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol
    if tname == "FFT" and dims[0] in (640, 768):
        # the long dim's passes (forward and backward) run a specialised kernel
        assert sum(",%d,c2c" % dims[0] in k for k in meta.kernels) == 2, meta.kernels


def test_mixed_workload_full_size(pf, gpu):
    """The mixed workload at its full size: (RFFT, FFT, R2R(REDFT10)) on a
    Float32 384x384x256 field, inverse through (BRFFT, BFFT, REDFT01), vs the
    FP64 oracle; every pass on the specialised kernels."""
    from oracle import pipeline as OP
    from oracle import plan as OM
    T = pf.Transforms
    dims = (384, 384, 256)
    trs = (T.RFFT(), T.FFT(), T.R2R(T.REDFT10))
    plan = pf.PencilFFTPlan(dims, trs, (1, 1), None, T.Float32)
    ks = plan.kernels(0) + plan.kernels(2)
    print(ks)
    assert all(k.startswith("k_fast") for k in ks), ks
    u = pf.allocate_input(plan)
    # the workload's data: 32 Fourier(x, y) x Chebyshev(z) modes + 1e-3 noise
    u.logical().copy_(pf.workloads.fourier_chebyshev_field(dims, n_modes=32, noise=1e-3,
                                                           seed=7))
    v = plan * u
    w = plan.ldiv(v)
    torch.cuda.synchronize()
    x = u.parent.cpu().numpy().astype(np.float64).reshape(-1)
    err_b = OP.rel_l2(w.parent.cpu().numpy().reshape(-1), x)
    meta = OM.PlanMeta(dims, ("RFFT", "FFT", "R2R_REDFT10"), "Float64")
    ref = OP.forward(meta, OP.from_image(x, dims), workers=16)
    err_f = OP.rel_l2(v.parent.cpu().numpy().reshape(-1), OP.memory_image(ref))
    print("mixed 384x384x256: fwd %.2e rt %.2e" % (err_f, err_b))
    tol = 5e-7 * np.log2(float(np.prod(dims)))
    assert err_f <= tol and err_b <= tol


def test_cpu_case_64_f64_round_trip(pf, gpu):
    """The reference's CPU-runnable case (64^3 Float64, uniform [0,1) with
    seed 42, RFFT then ldiv!): spectrum against numpy's rfftn, round trip."""
    T = pf.Transforms
    dims = (64, 64, 64)
    plan = pf.PencilFFTPlan(dims, T.RFFT(), (1, 1), None, T.Float64)
    x = np.random.default_rng(42).uniform(0.0, 1.0, size=dims)
    u = pf.allocate_input(plan)
    u.logical().copy_(torch.from_numpy(x).to("cuda"))
    v = plan * u
    # numpy's rfftn halves the last axis; the plan halves logical axis 0
    ref = np.fft.rfftn(x.transpose(2, 1, 0)).transpose(2, 1, 0)
    got = v.logical().cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref) * np.log2(x.size)
    w = plan.ldiv(v)
    assert np.max(np.abs(w.logical().cpu().numpy() - x)) <= 1e-14


@pytest.mark.parametrize("dims,real_type,permute", [
    ((256, 512, 128), 0, True), ((512, 128, 96), 0, True), ((128, 256, 64), 1, True),
    ((128, 512, 32), 1, True)],
    ids=["256x512x128-F32", "512x128x96-F32", "128x256x64-F64", "128x512x32-F64"])
def test_prefetching_tile_loads(pf, gpu, dims, real_type, permute):
    """The prefetching transposed-load c2r kernels (persistent grid, several
    tiles per workgroup, the next tile's loads in flight), forced one at a
    time, against the oracle."""
    T = pf.Transforms
    L = pf._lib
    tr = T.RFFT()
    seen = {}
    for k in range(16):
        plan = pf.PencilFFTPlan(dims, tr, (1, 1), None, real_type, permute_dims=permute,
                                flags=L.FLAG_CANDIDATE(k))
        for name in plan.kernels(0) + plan.kernels(2):
            if ",PF" in name and name not in seen:
                seen[name] = k
        del plan
    assert seen, "no prefetching kernel among the candidates"
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    for name, k in seen.items():
        for rev in (0, L.FLAG_REVERSE_TILES, L.FLAG_REVERSE_TILES | L.FLAG_STREAM_STORES):
            err_f, err_b, meta = run_case(pf, dims, (tr,), real_type, permute=permute, seed=k,
                                          flags=L.FLAG_CANDIDATE(k) | rev)
            print(name, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
            # the kernel itself, whatever direction / streaming bits it runs with
            base = lambda s: s.replace(",rev", "").replace(",nt", "")
            assert any(base(x) == base(name) for x in meta.kernels), (name, meta.kernels)
            assert err_f <= tol and err_b <= tol, (name, rev, err_f, err_b)


@pytest.mark.parametrize("tname", ["RFFT", "FFT"])
@pytest.mark.parametrize("dims", [(384, 96, 64), (192, 384, 32)], ids=lambda d: "x".join(map(str, d)))
def test_fast_mixed_radix_f64(pf, gpu, dims, tname):
    """Float64 lines of 96, 192 and 384 points (also as the half-length FFTs
    of 384 reals) on the specialised kernels, at the Float64 bound."""
    T = pf.Transforms
    err_f, err_b, meta = run_case(pf, dims, (getattr(T, tname)(),), 1, seed=6)
    print(dims, tname, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= 1e-13 and err_b <= 1e-13
    assert all(k.startswith("k_fast<f64") for k in meta.kernels), meta.kernels


@pytest.mark.parametrize("permute", [True, False], ids=["perm", "noperm"])
@pytest.mark.parametrize("dims", [(2048, 32, 16), (16, 2048, 32), (4096, 16, 8)],
                         ids=lambda d: "x".join(map(str, d)))
def test_fast_2048(pf, gpu, dims, permute):
    """2048-point FP32 c2c lines, and the r2c / c2r of 4096 reals, against
    the oracle; permuted, their passes run specialised kernels."""
    T = pf.Transforms
    tr = T.RFFT() if dims[0] == 4096 else T.FFT()
    err_f, err_b, meta = run_case(pf, dims, (tr,), 0, permute=permute, seed=8)
    tol = 5e-7 * np.log2(int(np.prod(dims)))
    print(dims, permute, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol
    if permute:
        kind = ",2048,r2c" if dims[0] == 4096 else ",2048,c2c"
        assert sum(kind in k for k in meta.kernels) == (1 if dims[0] == 4096 else 2), meta.kernels
