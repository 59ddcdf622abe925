"""Every specialised kernel variant the autotuner may pick, forced one at a
time (PFFT_FLAG_CANDIDATE(k)), against the FP64 oracle.  The autotuner only
chooses among these by speed, so each of them must be right on its own."""
import numpy as np
import pytest

from tests.test_gpu_parity import run_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_CANDIDATES = 24

# (dims, transforms, real type, precision flags): each shape puts its passes
# on line->line, line->tile and tile->line kernels of one length family
CASES = [
    ((512, 64, 32), "FFT", 0, 0),
    ((512, 32, 48), "RFFT", 0, 0),
    ((256, 64, 32), "RFFT", 0, 0),
    ((1024, 32, 16), "FFT", 0, 0),
    ((1024, 32, 16), "FFT", 0, 4),
    ((1024, 32, 16), "FFT", 0, 8),
    ((32, 1024, 16), "FFT", 0, 4),
    ((384, 48, 32), "FFT", 0, 0),
    ((256, 32, 16), "RFFT", 1, 0),
    ((512, 16, 16), "FFT", 1, 0),
    ((64, 32, 256), "MIXED", 0, 0),
    ((64, 64, 32), "RFFT", 1, 0),
    ((32, 64, 256), "MIXED", 1, 0),
    ((128, 32, 64), "MIXED", 1, 0),
]


def _trs(T, name):
    if name == "MIXED":
        return (T.RFFT(), T.FFT(), T.R2R(T.REDFT10))
    return (getattr(T, name)(),)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s-f%d" % (
    "x".join(map(str, c[0])), c[1], "F64" if c[2] else "F32", c[3]))
def test_every_candidate(pf, gpu, case):
    dims, name, real_type, flags = case
    T = pf.Transforms
    L = pf._lib
    seen = set()
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    for k in range(MAX_CANDIDATES):
        err_f, err_b, meta = run_case(pf, dims, _trs(T, name), real_type, seed=k,
                                      flags=flags | L.FLAG_CANDIDATE(k))
        new = [x for x in meta.kernels if x not in seen]
        seen.update(meta.kernels)
        print(k, "fwd %.2e rt %.2e" % (err_f, err_b), new, flush=True)
        assert err_f <= tol and err_b <= tol, (k, meta.kernels, err_f, err_b)
    if flags:
        assert any("mfma" in x for x in seen), seen
    print(len(seen), "kernels:", sorted(seen))


REVERSED = [CASES[1], CASES[2], CASES[3], CASES[8], CASES[10]]


@pytest.mark.parametrize("case", REVERSED, ids=lambda c: "%s-%s-%s-f%d" % (
    "x".join(map(str, c[0])), c[1], "F64" if c[2] else "F32", c[3]))
def test_every_candidate_tiles_reversed(pf, gpu, case):
    """The same candidates walking their tiles in descending order
    (PFFT_FLAG_REVERSE_TILES), the direction the autotuner may pick per pass;
    every other candidate also stores with the nontemporal hint
    (PFFT_FLAG_STREAM_STORES), the other tuned bit."""
    dims, name, real_type, flags = case
    T = pf.Transforms
    L = pf._lib
    n = int(np.prod(dims))
    tol = 1e-13 if real_type == 1 else 5e-7 * np.log2(n)
    seen = set()
    for k in range(MAX_CANDIDATES):
        err_f, err_b, meta = run_case(pf, dims, _trs(T, name), real_type, seed=k,
                                      flags=flags | L.FLAG_REVERSE_TILES | L.FLAG_CANDIDATE(k)
                                      | (L.FLAG_STREAM_STORES if k % 2 else 0))
        seen.update(meta.kernels)
        assert err_f <= tol and err_b <= tol, (k, meta.kernels, err_f, err_b)
    assert any(",rev" in x for x in seen), seen
    assert any(",nt" in x for x in seen), seen


def test_xlane_bit_identical(pf, gpu):
    """The 256-, 512- and 1024-point kernels whose last stage exchange runs across
    lanes (DPP row rotation and the gfx950 row-swap instructions, `xlane` in
    the name) do the Stockham form's arithmetic in another data movement:
    every forced candidate of each plan below gives the bit-identical output,
    the first one matches numpy's float64 transform, and the candidates cover
    the cross-lane form of every 256/512-point pass shape the plans use."""
    T = pf.Transforms
    L = pf._lib
    seen = set()
    # 1024-point lines also have two-wavefront variants (another radix
    # schedule, equal to rounding only): those plans are checked per candidate
    for dims, name in (((512, 64, 32), "FFT"), ((32, 64, 512), "FFT"), ((512, 64, 32), "RFFT"),
                       ((256, 64, 32), "FFT"), ((32, 32, 256), "FFT"), ((1024, 32, 16), "FFT"),
                       ((16, 32, 1024), "FFT")):
        exact = 1024 not in dims
        rng = np.random.default_rng(5)
        if name == "RFFT":
            x = rng.standard_normal(dims).astype(np.float32)
            want = np.fft.rfftn(x.astype(np.float64), axes=(2, 1, 0))   # r2c along dim 1
        else:
            x = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(np.complex64)
            want = np.fft.fftn(x.astype(np.complex128))
        ref = None
        for k in range(48):
            plan = pf.PencilFFTPlan(dims, getattr(T, name)(), (1, 1), None, T.Float32,
                                    flags=L.FLAG_CANDIDATE(k))
            u = pf.allocate_input(plan)
            u.logical().copy_(torch.from_numpy(x))
            v = plan * u
            torch.cuda.synchronize()
            out = v.parent.cpu().numpy()
            seen.update(plan.kernels(0) + plan.kernels(2))
            if ref is None or not exact:
                ref = out
                got = v.logical().cpu().numpy()
                err = np.linalg.norm(got - want) / np.linalg.norm(want)
                assert err <= 5e-7 * np.log2(x.size), (dims, name, k, err, plan.kernels(0))
            else:
                assert np.array_equal(out, ref), (dims, name, k, plan.kernels(0))
    xl = sorted(s for s in seen if "xlane" in s)
    print(xl)
    assert any(",1024," in s for s in xl), xl
    for m in (",512,", ",256,"):
        for kind in ("c2c", "r2c", "c2r"):
            for form in ("line->line", "line->tile", "tile->line"):
                if any(m + kind + "," + form in s for s in seen):
                    assert any(m + kind + "," + form in s for s in xl), (m, kind, form, xl)


def test_mixed_radix_row_chunks(pf, gpu):
    """Row-chunked transposed stores (`S=` in the name) of the mixed-radix
    lengths 96-384: every forced candidate of each plan within the Float32
    bound of numpy's float64 transform and within rounding of the first
    candidate (the chunking changes no arithmetic, but another template
    instance may contract multiply-adds differently), and the chunked kernels
    of the c2c and r2c passes are among the candidates."""
    T = pf.Transforms
    L = pf._lib
    seen = set()
    for dims, name in (((384, 48, 32), "FFT"), ((320, 48, 32), "FFT"), ((640, 32, 16), "RFFT"),
                       ((384, 32, 16), "RFFT"), ((192, 32, 16), "FFT")):
        rng = np.random.default_rng(9)
        if name == "RFFT":
            x = rng.standard_normal(dims).astype(np.float32)
            want = np.fft.rfftn(x.astype(np.float64), axes=(2, 1, 0))   # r2c along dim 1
        else:
            x = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(np.complex64)
            want = np.fft.fftn(x.astype(np.complex128))
        tol = 5e-7 * np.log2(x.size)
        ref = None
        for k in range(40):
            plan = pf.PencilFFTPlan(dims, getattr(T, name)(), (1, 1), None, T.Float32,
                                    flags=L.FLAG_CANDIDATE(k))
            u = pf.allocate_input(plan)
            u.logical().copy_(torch.from_numpy(x))
            v = plan * u
            torch.cuda.synchronize()
            got = v.logical().cpu().numpy()
            seen.update(plan.kernels(0) + plan.kernels(2))
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert err <= tol, (dims, name, k, err, plan.kernels(0))
            if ref is None:
                ref = got
            else:
                d = np.linalg.norm(got - ref) / np.linalg.norm(ref)
                assert d <= 1e-6, (dims, name, k, d, plan.kernels(0))
    chunked = sorted(s for s in seen if ",S=" in s and any(",%d," % m in s for m in (192, 320, 384)))
    print(chunked)
    for m in (",320,", ",384,"):
        assert any(m + "c2c" in s for s in chunked), (m, chunked)
    assert any("r2c" in s for s in chunked), chunked
