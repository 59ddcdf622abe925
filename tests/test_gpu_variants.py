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
