"""The nine R2R kinds of the reference (src/Transforms/r2r.jl:5-15, planned at
:86-93) on the GPU against the FP64 oracle, at the lengths the specialised
passes cover: real data ((R2R(kind),) x 3: every pass an R2R of real lines)
and complex data ((FFT, R2R(kind), R2R(kind))).  REDFT10 / REDFT01 /
RODFT10 / RODFT01 / REDFT11 / RODFT11 / DHT run on the specialised kernels
(pre/post-twiddled FFTs; two real lines per complex FFT), REDFT00 / RODFT00
on the generic pass at these lengths (extended lines of 510 / 514 points); at
n = 2^k + 1 / 2^k - 1 they have specialised kernels too (the last tests).
Each case records which kernels ran.  Every specialised candidate is also
forced once."""
import numpy as np
import pytest

from tests.test_gpu_parity import run_case

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KINDS = ["REDFT00", "REDFT01", "REDFT10", "REDFT11", "RODFT00", "RODFT01", "RODFT10",
         "RODFT11", "DHT"]
FAST = {"REDFT01", "REDFT10", "REDFT11", "RODFT01", "RODFT10", "RODFT11", "DHT"}


def _trs(T, kind, data):
    k = getattr(T, kind)
    if data == "real":
        return (T.R2R(k),)
    return (T.FFT(), T.R2R(k), T.R2R(k))


@pytest.mark.parametrize("real_type", [0, 1], ids=["F32", "F64"])
@pytest.mark.parametrize("data", ["real", "complex"])
@pytest.mark.parametrize("kind", KINDS)
def test_r2r_kinds(pf, gpu, kind, data, real_type):
    T = pf.Transforms
    dims = (256, 64, 128) if real_type == 0 else (256, 64, 64)
    err_f, err_b, meta = run_case(pf, dims, _trs(T, kind, data), real_type, seed=11)
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    print(kind, data, real_type, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol
    if kind in FAST:
        # every R2R pass on the specialised kernels
        assert all(k.startswith("k_fast") for k in meta.kernels), meta.kernels


def test_r2r_real_chebyshev_512(pf, gpu):
    """A pure Chebyshev plan, (R2R(REDFT10),) x 3 on a Float32 512^3 field, runs
    only specialised passes; its error against the oracle and round trip."""
    from oracle import pipeline as OP
    from oracle import plan as OM
    T = pf.Transforms
    dims = (512, 512, 512)
    plan = pf.PencilFFTPlan(dims, T.R2R(T.REDFT10), (1, 1), None, T.Float32)
    ks = plan.kernels(0) + plan.kernels(2)
    print(ks)
    assert all(k.startswith("k_fast") for k in ks), ks
    u = pf.allocate_input(plan)
    u.parent.uniform_(-1, 1, generator=torch.Generator(device="cuda").manual_seed(3))
    v = plan * u
    w = plan.ldiv(v)
    torch.cuda.synchronize()
    x = u.parent.cpu().numpy().astype(np.float64).reshape(-1)
    err_b = OP.rel_l2(w.parent.cpu().numpy().reshape(-1), x)
    meta = OM.PlanMeta(dims, ("R2R_REDFT10",), "Float64")
    ref = OP.forward(meta, OP.from_image(x, dims), workers=16)
    err_f = OP.rel_l2(v.parent.cpu().numpy().reshape(-1), OP.memory_image(ref))
    print("512^3 REDFT10 x3: fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= 4.5e-6 and err_b <= 4.5e-6


MAX_CANDIDATES = 16


@pytest.mark.parametrize("real_type", [0, 1], ids=["F32", "F64"])
@pytest.mark.parametrize("data", ["real", "complex"])
@pytest.mark.parametrize("kind", sorted(FAST))
def test_r2r_every_candidate(pf, gpu, kind, data, real_type):
    T = pf.Transforms
    L = pf._lib
    dims = (128, 64, 64) if real_type == 0 else (64, 64, 64)
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    seen = set()
    for k in range(MAX_CANDIDATES):
        rev = L.FLAG_REVERSE_TILES | L.FLAG_STREAM_STORES if k % 2 else 0
        err_f, err_b, meta = run_case(pf, dims, _trs(T, kind, data), real_type, seed=k,
                                      flags=L.FLAG_CANDIDATE(k) | rev)
        seen.update(meta.kernels)
        assert err_f <= tol and err_b <= tol, (k, meta.kernels, err_f, err_b)
    print(kind, data, real_type, sorted(seen))
    tag = "r," if data == "real" else "c,"
    assert any(tag in s and ("dct" in s or "dst" in s or "dht" in s) for s in seen), seen


@pytest.mark.parametrize("kind", ["REDFT10", "REDFT01"])
def test_dct_row_chunks(pf, gpu, kind):
    """The row-chunked transposed stores of the 256-point DCT-II / DCT-III
    kernels of complex data (`S=2`), forced among the candidates, against the
    oracle."""
    T = pf.Transforms
    L = pf._lib
    dims = (32, 256, 64)   # FFT along 32, the DCT of complex data along 256 and 64
    tol = 5e-7 * np.log2(int(np.prod(dims)))
    seen = set()
    for k in range(24):
        err_f, err_b, meta = run_case(pf, dims, _trs(T, kind, "complex"), 0, seed=k,
                                      flags=L.FLAG_CANDIDATE(k))
        seen.update(meta.kernels)
        assert err_f <= tol and err_b <= tol, (k, meta.kernels, err_f, err_b)
    chunked = [s for s in seen if ",256,dct" in s and ",S=2" in s]
    print(sorted(chunked))
    assert chunked, sorted(seen)


# DCT-I / DST-I lengths with a specialised kernel: n = M/2 + 1 (REDFT00) and
# n = M/2 - 1 (RODFT00) for the extended lines' FFT length M
EXT_LENGTHS = {0: {"REDFT00": [129, 257, 513], "RODFT00": [127, 255, 511]},
               1: {"REDFT00": [129, 257], "RODFT00": [127, 255]}}
EXT_CASES = [(kind, n, data, rt) for rt in (0, 1) for kind in ("REDFT00", "RODFT00")
             for n in EXT_LENGTHS[rt][kind] for data in ("complex", "real")]


def _ext_case(T, kind, n, data):
    k = T.R2R(getattr(T, kind))
    if data == "real":   # the last dim's real lines as pairs
        return (64, 64, n), (T.R2R(T.REDFT10), T.R2R(T.REDFT10), k)
    return (32, 32, n), (T.FFT(), T.FFT(), k)


@pytest.mark.parametrize("kind,n,data,real_type", EXT_CASES,
                         ids=["%s-%d-%s-%s" % (k, n, d, "F32" if r == 0 else "F64")
                              for k, n, d, r in EXT_CASES])
def test_dct1_dst1_specialised(pf, gpu, kind, n, data, real_type):
    """REDFT00 / RODFT00 (src/Transforms/r2r.jl:5-15) along the last dim at
    the lengths of the specialised kernels (the M-point FFT of the even / odd
    extension, M a power of two), complex data and real line pairs, against
    the oracle.  The pass has one transposed side (the load forward, the
    store backward); both directions run the specialised kernel."""
    T = pf.Transforms
    dims, trs = _ext_case(T, kind, n, data)
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(int(np.prod(dims)))
    err_f, err_b, meta = run_case(pf, dims, trs, real_type, seed=n)
    tag = ",dct1" if kind == "REDFT00" else ",dst1"
    ext = [s for s in meta.kernels if tag in s]
    print(kind, n, data, real_type, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert len(ext) == 2, meta.kernels
    assert any("tile->line" in s for s in ext) and any("line->tile" in s for s in ext), ext
    assert err_f <= tol and err_b <= tol


@pytest.mark.parametrize("real_type", [0, 1], ids=["F32", "F64"])
@pytest.mark.parametrize("data", ["real", "complex"])
@pytest.mark.parametrize("kind", ["REDFT00", "RODFT00"])
def test_dct1_dst1_every_candidate(pf, gpu, kind, data, real_type):
    """Every specialised DCT-I / DST-I candidate forced once (both tile
    orders, reversed tiles, streaming stores)."""
    T = pf.Transforms
    L = pf._lib
    n = 257 if kind == "REDFT00" else 255
    dims, trs = _ext_case(T, kind, n, data)
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(int(np.prod(dims)))
    seen = set()
    for k in range(8):
        rev = L.FLAG_REVERSE_TILES | L.FLAG_STREAM_STORES if k % 2 else 0
        err_f, err_b, meta = run_case(pf, dims, trs, real_type, seed=k,
                                      flags=L.FLAG_CANDIDATE(k) | rev)
        seen.update(meta.kernels)
        assert err_f <= tol and err_b <= tol, (k, meta.kernels, err_f, err_b)
    tag = ",dct1" if kind == "REDFT00" else ",dst1"
    print(kind, data, real_type, sorted(s for s in seen if tag in s))
    assert any(tag in s for s in seen), seen


@pytest.mark.parametrize("data", ["real", "complex"])
@pytest.mark.parametrize("kind", ["REDFT10", "REDFT01", "REDFT11"])
def test_dct_xlane_candidates(pf, gpu, kind, data):
    """The 256- and 512-point DCT-II / III / IV kernels with the cross-lane
    last exchange, forced among the candidates, against the oracle."""
    T = pf.Transforms
    L = pf._lib
    dims = (512, 256, 32) if data == "real" else (32, 512, 256)
    tol = 5e-7 * np.log2(int(np.prod(dims)))
    seen = set()
    for k in range(24):
        err_f, err_b, meta = run_case(pf, dims, _trs(T, kind, data), 0, seed=k,
                                      flags=L.FLAG_CANDIDATE(k))
        seen.update(meta.kernels)
        assert err_f <= tol and err_b <= tol, (k, meta.kernels, err_f, err_b)
    xl = sorted(s for s in seen if ",xlane" in s and ("dct" in s))
    print(kind, data, xl)
    assert xl, sorted(seen)


@pytest.mark.parametrize("permute,extra", [(False, ()), (True, (3,)), (False, (2,))],
                         ids=["noperm", "extra3", "noperm-extra2"])
@pytest.mark.parametrize("kind,n", [("REDFT00", 257), ("RODFT00", 255), ("REDFT11", 256),
                                    ("RODFT11", 256)])
def test_new_kinds_layouts(pf, gpu, kind, n, permute, extra):
    """DCT-I / DST-I / DCT-IV / DST-IV passes under the layout options of
    transforms.jl:414-462 (permute_dims=false) and with extra dims
    (plans.jl:164-182), complex data, against the oracle."""
    T = pf.Transforms
    dims = (32, 32, n)
    trs = (T.FFT(), T.FFT(), T.R2R(getattr(T, kind)))
    tol = 5e-7 * np.log2(int(np.prod(dims) * int(np.prod(extra or (1,)))))
    err_f, err_b, meta = run_case(pf, dims, trs, 0, permute=permute, extra=extra, seed=n)
    print(kind, n, permute, extra, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol


@pytest.mark.parametrize("data", ["real", "complex"])
@pytest.mark.parametrize("dims", [(96, 192, 64), (384, 160, 32), (64, 320, 96)],
                         ids=lambda d: "x".join(map(str, d)))
def test_dct_mixed_lengths(pf, gpu, dims, data):
    """REDFT10 / REDFT01 at the mixed-radix lengths 96, 160, 192, 320 and 384
    (the Chebyshev pair on 3*2^k and 5*2^k points) on the specialised
    kernels, against the oracle."""
    T = pf.Transforms
    err_f, err_b, meta = run_case(pf, dims, _trs(T, "REDFT10", data), 0, seed=3)
    tol = 5e-7 * np.log2(int(np.prod(dims)))
    print(dims, data, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= tol and err_b <= tol
    mixed = [k for k in meta.kernels if ",dct" in k and any(",%d," % m in k for m in (96, 160, 192, 320, 384))]
    assert mixed, meta.kernels


@pytest.mark.parametrize("data", ["real", "complex"])
def test_dct_f64_128_512(pf, gpu, data):
    """Float64 REDFT10 / REDFT01 on 512- and 128-point lines (the specialised
    DCT-II / DCT-III kernels), against the oracle at the Float64 bound."""
    T = pf.Transforms
    dims = (512, 128, 32) if data == "real" else (32, 512, 128)
    err_f, err_b, meta = run_case(pf, dims, _trs(T, "REDFT10", data), 1, seed=9)
    print(dims, data, meta.kernels, "fwd %.2e rt %.2e" % (err_f, err_b))
    assert err_f <= 1e-13 and err_b <= 1e-13
    assert any(k.startswith("k_fast<f64,512,dct") for k in meta.kernels), meta.kernels
    assert any(k.startswith("k_fast<f64,128,dct") for k in meta.kernels), meta.kernels
