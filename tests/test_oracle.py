"""Pins of the CPU oracle (oracle/).

The reference cannot run here (no Julia, no FFTW), so each oracle function is
checked against what fixes it independently of its own formula:
  * metadata: the reference's doctest known answers (Transforms.jl:100-106,
    121-137, 162-187, 222-259, 273-283; global_params.jl:31-40) and the
    assertions of test/transforms.jl:19-143;
  * 1D transforms: brute-force long-double sums vs pocketfft (two independent
    evaluations), the unnormalised inverse pairs stated in the FFTW manual
    (REDFT10/REDFT01 -> 2n, REDFT00 -> 2(n-1), RODFT00 -> 2(n+1), DHT -> n,
    FFT/BFFT -> n, RFFT/BRFFT -> n), textbook values (DFT of [1,2,3,4],
    impulses, constants), Parseval;
  * pipeline: the closed-form index map of the permutation chain, and
    separability (the N-D result equals numpy's n-dimensional transform).
"""
import numpy as np
import pytest

from oracle import pipeline as OP
from oracle import plan as OM
from oracle import transforms1d as T1

KINDS = ["R2R_" + k for k in OM.R2R_KINDS]


# ---------------------------------------------------------------------------
# metadata known answers
def test_doctest_binv():
    assert OM.binv("FFT", 42) == "BFFT"                 # Transforms.jl:101-102
    assert OM.binv("BRFFT_odd", 42) == "RFFT"           # BRFFT(9), :104-105
    assert OM.binv("RFFT", 10) == "BRFFT_even"          # transforms.jl:25
    assert OM.binv("RFFT", 11) == "BRFFT_odd"           # transforms.jl:26
    assert OM.binv("FFT!", 4) == "BFFT!" and OM.binv("BFFT!", 4) == "FFT!"
    assert OM.binv("R2R_REDFT01", 42) == "R2R_REDFT10"  # transforms.jl:117-119
    assert OM.binv("NoTransform", 42) == "NoTransform"
    for k in OM.R2R_KINDS:
        assert OM.binv(OM.binv("R2R_" + k, 5), 5) == "R2R_" + k


def test_doctest_is_inplace():
    # Transforms.jl:121-137
    assert OM.is_inplace("RFFT") is False
    assert OM.is_inplace("NoTransform!") is True
    assert OM.is_inplace_list(["FFT!", "R2R_REDFT01!"]) is True
    assert OM.is_inplace_list(["FFT", "R2R_REDFT01"]) is False
    assert OM.is_inplace_list(["FFT", "R2R_REDFT01!"]) is None
    # transforms.jl:131-138
    assert OM.is_inplace_list(["FFT", "FFT", "FFT!"]) is None
    assert OM.is_inplace_list(["FFT!", "FFT", "FFT!"]) is None
    assert OM.is_inplace_list(["FFT!", "FFT!", "FFT!"]) is True
    with pytest.raises(ValueError, match="cannot combine in-place and out-of-place"):
        OM.PlanMeta((16, 12, 6), ("FFT", "FFT!", "FFT!"))


def test_doctest_scale_factor():
    # Transforms.jl:162-187
    C = (3, 4, 5)
    assert OM.scale_factor_array("FFT", C, (1, 2, 3)) == 60
    assert OM.scale_factor_array("BFFT", C, (1, 2, 3)) == 60
    assert OM.scale_factor_array("BFFT", C, (2, 3)) == 20
    assert OM.scale_factor_array("RFFT", C, (2,)) == 4
    assert OM.scale_factor_array("RFFT", C, (2, 3)) == 20
    assert OM.scale_factor_array("BRFFT_even", C, (1, 2, 3)) == 96
    assert OM.scale_factor_array("BRFFT_odd", C, (1, 2, 3)) == 108
    assert OM.scale_factor_array("FFT", (3, 4), (1, 2)) == 12    # transforms.jl:141


def test_doctest_eltypes():
    # Transforms.jl:222-259
    assert OM.eltype_input("FFT", "Float32") == "ComplexF32"
    assert OM.eltype_input("RFFT", "Float64") == "Float64"
    assert OM.eltype_input("R2R_REDFT01", "Float64") is None
    assert OM.eltype_input("NoTransform", "Float64") is None
    assert OM.eltype_output("NoTransform", "Float32") == "Float32"
    assert OM.eltype_output("RFFT", "Float64") == "ComplexF64"
    assert OM.eltype_output("BRFFT_even", "ComplexF32") == "Float32"
    with pytest.raises(ValueError, match="invalid input data type for "
                                         "PencilFFTs.Transforms.FFT: Float64"):
        OM.eltype_output("FFT", "Float64")
    # transforms.jl:349-365
    assert OM.input_data_type("Float32", ("NoTransform", "FFT")) == "ComplexF32"
    assert OM.input_data_type("Float32", ("NoTransform", "NoTransform")) == "Float32"
    assert OM.input_data_type("Float64", ("RFFT", "FFT", "FFT")) == "Float64"


def test_doctest_expand_dims():
    # Transforms.jl:273-283
    assert OM.expand_dims("RFFT", 3) == ("RFFT", "FFT", "FFT")
    assert OM.expand_dims("BRFFT_even", 3) == ("BFFT", "BFFT", "BRFFT_even")
    assert OM.expand_dims("NoTransform", 2) == ("NoTransform", "NoTransform")


def test_doctest_global_params():
    # global_params.jl:31-40 and transforms.jl:20-37
    m = OM.PlanMeta((64, 32, 128), ("RFFT", "FFT", "FFT"))
    assert m.size_out == (33, 32, 128)
    assert m.eltype_in == "Float64"
    size_in = (16, 12, 6)
    trs = ("RFFT", "FFT", "FFT")
    binvs = tuple(OM.binv(t, n) for t, n in zip(trs, size_in))
    assert binvs == ("BRFFT_even", "BFFT", "BFFT")
    size_out = tuple(OM.length_output(t, n) for t, n in zip(trs, size_in))
    assert size_out == (9, 12, 6)
    assert tuple(OM.length_output(t, n) for t, n in zip(binvs, size_out)) == size_in


def test_unsupported_r2r_kinds():
    # transforms.jl:108-111, r2r.jl:42-45
    for k in ("R2HC", "HC2R"):
        with pytest.raises(ValueError, match="unsupported r2r transform kind: " + k):
            OM.check("R2R_" + k)


def test_stage_chain():
    # plans.jl:503-521
    assert [OM.stage_perm(3, n) for n in range(3)] == [(0, 1, 2), (1, 0, 2), (2, 1, 0)]
    assert [OM.stage_perm(4, n) for n in range(4)] == [(0, 1, 2, 3), (1, 0, 2, 3),
                                                       (2, 0, 1, 3), (3, 2, 1, 0)]
    assert [OM.stage_perm(2, n) for n in range(2)] == [(0, 1), (1, 0)]
    assert [OM.stage_perm(3, n, False) for n in range(3)] == [(0, 1, 2)] * 3
    m = OM.PlanMeta((512, 512, 512), ("RFFT",), "Float32")
    assert m.mem_dims("in") == (512, 512, 512) and m.mem_dims("out") == (512, 512, 257)
    assert m.scale == 512.0 ** 3
    assert [s["mem_axis"] for s in m.stages] == [0, 0, 0]


def test_scale_factors_of_configs():
    # the workloads' normalisations (closed forms)
    assert OM.PlanMeta((64, 64, 64), ("RFFT",)).scale == 262144
    assert OM.PlanMeta((256,) * 3, ("RFFT",)).scale == 16777216
    assert OM.PlanMeta((1024,) * 3, ("FFT",)).scale == 1073741824
    m = OM.PlanMeta((384, 384, 256), ("RFFT", "FFT", "R2R_REDFT10"), "Float32")
    assert m.scale == 384 * 384 * 512 == 75497472
    assert m.size_out == (193, 384, 256)
    assert m.eltype_out == "ComplexF32"
    assert [s["tr_bwd"] for s in m.stages] == ["BRFFT_even", "BFFT", "R2R_REDFT01"]
    assert OM.PlanMeta((5, 6), ("R2R_REDFT00", "R2R_RODFT00")).scale == 2 * 4 * 2 * 7


# ---------------------------------------------------------------------------
# 1D transforms
SIZES = list(range(1, 18)) + [24, 32, 41]


@pytest.mark.parametrize("n", SIZES)
def test_direct_vs_fast(n):
    rng = np.random.default_rng(n)
    xr = rng.standard_normal((n, 3))
    xc = xr + 1j * rng.standard_normal((n, 3))
    for tr, x in [("FFT", xc), ("BFFT", xc), ("RFFT", xr)] + [(k, xr) for k in KINDS] + \
                 [(k, xc) for k in KINDS]:
        if tr == "R2R_REDFT00" and n < 2:
            continue
        a = T1.direct(tr, x, axis=0)
        b = T1.fast(tr, x, axis=0)
        assert np.linalg.norm(a - b) <= 1e-13 * max(np.linalg.norm(a), 1.0), (tr, n)
    h = n // 2 + 1
    X = rng.standard_normal((h, 2)) + 1j * rng.standard_normal((h, 2))
    brf = "BRFFT_even" if n % 2 == 0 else "BRFFT_odd"
    a = T1.direct(brf, X, axis=0)
    b = T1.fast(brf, X, axis=0)
    assert a.shape[0] == n
    assert np.linalg.norm(a - b) <= 1e-13 * max(np.linalg.norm(a), 1.0), n


def test_textbook_values():
    x = np.array([1.0, 2.0, 3.0, 4.0])
    want = np.array([10, -2 + 2j, -2, -2 - 2j])
    for f in (T1.direct, T1.fast):
        assert np.allclose(f("FFT", x.astype(complex)), want, atol=1e-14)
        assert np.allclose(f("RFFT", x), want[:3], atol=1e-14)
    n = 7
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1
        k = np.arange(n)
        assert np.allclose(T1.direct("FFT", e), np.exp(-2j * np.pi * j * k / n), atol=1e-15)
        assert np.allclose(T1.direct("BFFT", e), np.exp(2j * np.pi * j * k / n), atol=1e-15)
    # DCT-II of a constant: 2n at k = 0, zero elsewhere
    c = T1.direct("R2R_REDFT10", np.ones(6))
    assert np.allclose(c, [12, 0, 0, 0, 0, 0], atol=1e-14)
    # c2r of a single mode m: 2 Re(a exp(2 pi i m j / n))
    for n in (8, 9):
        X = np.zeros(n // 2 + 1, complex)
        X[2] = 1 + 2j
        j = np.arange(n)
        want = 2 * np.real((1 + 2j) * np.exp(2j * np.pi * 2 * j / n))
        brf = "BRFFT_even" if n % 2 == 0 else "BRFFT_odd"
        assert np.allclose(T1.direct(brf, X), want, atol=1e-14)
        # imaginary parts of DC (and Nyquist) are ignored
        X2 = X.copy()
        X2[0] += 5j
        if n % 2 == 0:
            X2[-1] += 3j
        assert np.allclose(T1.direct(brf, X2), want, atol=1e-14)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 11, 12, 16])
def test_fftw_inverse_pairs(n):
    """FFTW manual 4.8.3-4.8.5: applying a kind and then its binv multiplies
    by the logical size (the scale factors of r2r.jl:96-101)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    for k in OM.R2R_KINDS:
        tr = "R2R_" + k
        if k == "REDFT00" and n < 2:
            continue
        y = T1.direct(OM.binv(tr, n), T1.direct(tr, x))
        assert np.allclose(y, OM.scale_factor_1d(tr, n) * x, atol=1e-12), k
    xc = x + 1j * rng.standard_normal(n)
    assert np.allclose(T1.direct("BFFT", T1.direct("FFT", xc)), n * xc, atol=1e-12)
    brf = OM.binv("RFFT", n)
    assert np.allclose(T1.direct(brf, T1.direct("RFFT", x)), n * x, atol=1e-12)
    # Parseval
    X = T1.direct("FFT", xc)
    assert np.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(xc) ** 2))


# ---------------------------------------------------------------------------
# pipeline
def test_index_map_closed_form():
    X, Y, Z = 5, 4, 3
    u = np.arange(X * Y * Z, dtype=float).reshape((X, Y, Z), order="F")
    m = OM.PlanMeta((X, Y, Z), ("NoTransform",))
    v = OP.memory_image(OP.forward(m, u))
    flat = OP.memory_image(u)
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                assert v[z + Z * (y + Y * x)] == flat[x + X * (y + Y * z)]
    m2 = OM.PlanMeta((X, Y, Z), ("NoTransform",), permute=False)
    assert np.array_equal(OP.memory_image(OP.forward(m2, u)), flat)


@pytest.mark.parametrize("dims", [(6, 5, 4), (8, 7), (5, 4, 3, 6)])
@pytest.mark.parametrize("permute", [True, False])
def test_separability(dims, permute):
    """The staged pipeline equals numpy's n-dimensional transforms."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal(dims)
    N = len(dims)
    m = OM.PlanMeta(dims, ("RFFT",), permute=permute)
    got = OP.from_memory(OP.forward(m, OP.to_memory(x, m.perm_in)), m.perm_out)
    want = np.fft.rfftn(x, axes=tuple(range(1, N)) + (0,))
    assert np.allclose(got, want, atol=1e-12)
    xc = x + 1j * rng.standard_normal(dims)
    m = OM.PlanMeta(dims, ("FFT",), permute=permute)
    got = OP.from_memory(OP.forward(m, OP.to_memory(xc, m.perm_in)), m.perm_out)
    assert np.allclose(got, np.fft.fftn(xc), atol=1e-12)


@pytest.mark.parametrize("trs", [("FFT",), ("RFFT",), ("BFFT",), ("NoTransform",)] +
                         [(k,) for k in KINDS] +
                         [("NoTransform", "RFFT", "FFT"), ("FFT", "NoTransform", "FFT")])
def test_round_trip(trs):
    dims = (16, 12, 6)
    m = OM.PlanMeta(dims, trs)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(dims)
    if m.eltype_in.startswith("Complex"):
        x = x + 1j * rng.standard_normal(dims)
    v = OP.forward(m, x)
    back = OP.backward(m, v)
    assert np.allclose(back, x, atol=1e-12)


def test_brfft_single_mode_closed_form():
    # test/brfft.jl:17-45 restated on the oracle
    for odd in (0, 1):
        dims_real = (12, 13, 16 + odd)
        c = np.zeros((12, 13, 9), complex)
        c[1, 3, 2] = 1 + 2j
        m = OM.PlanMeta((12, 13, 9), ("BRFFT_even" if not odd else "BRFFT_odd",))
        assert m.size_out == dims_real
        got = OP.from_memory(OP.forward(m, c), m.perm_out)
        x, y, z = np.meshgrid(np.arange(12), np.arange(13), np.arange(dims_real[2]),
                              indexing="ij")
        want = 2 * np.real((1 + 2j) * np.exp(2j * np.pi * (x / 12 + 3 * y / 13 +
                                                           2 * z / dims_real[2])))
        assert np.allclose(got, want, atol=1e-12)
